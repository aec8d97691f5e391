"""Writes the FP64-oracle fixtures of the scatter workloads at the sizes the
benchmark times (too slow for the oracle inside a test):

  cfg4_crystal_grid_0p25mm.npz   config 4 (2 layers of 32 x 32 bars, 32 x 32 x 1
                                 voxels), scatter points at the centres of
                                 10 x 32 x 10 cells (0.25 mm) of every bar,
                                 64-node power-law mu(E) tables: 12 seeded
                                 elements of the benchmark's detector rows
                                 (every 256th bar), with the oracle's own
                                 crystal bitmap.  Keys d, v (int32), value
                                 (float64), grid, rows.

Run from the repository root:  python tests/golden/make_scatter_golden.py
(about 3 minutes on 8 host threads).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-based-system-matrix-calculation-for-spect-pet_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from smg import params  # noqa: E402

GRID = (10, 32, 10)


def crystal_grid_fixture():
    import oracle

    p = params.config(4)
    o = oracle.Oracle(p)
    t0 = time.time()
    bm = o.bitmap_crystal()
    print(f"crystal bitmap of {p.n_det} bars: {time.time() - t0:.0f} s", flush=True)
    o.set_mu_tables(*params.synthetic_mu_tables(True))
    rows = np.arange(0, p.n_det, 256, dtype=np.int32)
    rng = np.random.default_rng(60)
    d = rows[rng.integers(0, len(rows), 12)].astype(np.int32)
    v = rng.integers(0, o.n_vox, 12).astype(np.int32)
    t0 = time.time()
    val = o.crystal_grid_scatter_pairs(bm, GRID, d, v)
    print(f"12 elements at {GRID}: {time.time() - t0:.0f} s", flush=True)
    np.savez(os.path.join(HERE, "cfg4_crystal_grid_0p25mm.npz"), d=d, v=v, value=val,
             grid=np.array(GRID, np.int32), rows=rows)


if __name__ == "__main__":
    crystal_grid_fixture()
