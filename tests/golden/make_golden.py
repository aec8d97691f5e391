"""Writes the committed fixtures of this directory.

  params_sha256.json   SHA-256 of the four Params_*.dat files of each
                       benchmark geometry (the seeded generators are pinned)
  cfg1_pe_fp64.npy     FP64 oracle PE matrix of config 1 (4 x 4 x 4 sub-cells),
                       64 x 1024 float64: a regression pin of the oracle and a
                       reference the GPU tests can read without recomputing

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-based-system-matrix-calculation-for-spect-pet_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from smg import params  # noqa: E402


def params_digests():
    out = {}
    for cfg in range(1, 6):
        p = params.config(cfg)
        with tempfile.TemporaryDirectory() as d:
            p.write(d)
            out[str(cfg)] = {name: hashlib.sha256(open(os.path.join(d, name), "rb").read()).hexdigest()
                             for name in sorted(params.FILE_NAMES.values())}
    return out


def main():
    import oracle

    with open(os.path.join(HERE, "params_sha256.json"), "w") as f:
        json.dump(params_digests(), f, indent=1, sort_keys=True)
    p = params.config(1)
    A = oracle.Oracle(p).pe_matrix(sub=(4, 4, 4))
    np.save(os.path.join(HERE, "cfg1_pe_fp64.npy"), A)


if __name__ == "__main__":
    main()
