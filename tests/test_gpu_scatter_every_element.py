"""Every element of the benchmark-size scatter matrices in FP32 against the
GPU's all-FP64 matrices of the same workloads, whose elements the sampled and
golden-fixture tests pin to the oracle (tests/test_gpu_scatter.py): within the
FP32 tolerance, and zero exactly where the FP64 element is (the same
energy-window, culling and bitmap decisions everywhere)."""
import numpy as np
import pytest

from conftest import TOL_FP32, rel_errors

pytestmark = pytest.mark.gpu


def _compare(got, ref, what):
    assert got.shape == ref.shape
    mx, med, zero_ok = rel_errors(got, ref)
    nz = int(np.count_nonzero(ref))
    print(f"{what}, every element: max rel {mx:.3e}, median {med:.3e}, {nz} of {ref.size} nonzero",
          flush=True)
    assert zero_ok, f"{what}: elements with |FP64| < 1e-30 differ"
    assert mx <= TOL_FP32, f"{what}: max relative error {mx:.3e}"
    assert nz > 0.05 * ref.size


def _tables():
    from smg import params
    return params.synthetic_mu_tables(True)


def test_crystal_scatter_bar_centres_2048_bars(smg_mod, params_mod):
    """The benchmark's bar-centre crystal scatter: 32x32x1 voxels x 2048 bars,
    2047 scatter points per row, 64-node mu(E) tables (2,097,152 elements)."""
    p = params_mod.config(4)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*p.meta["pe_subsamples"])
        ctx.set_scatter_subsamples(*p.meta["subsamples"])
        ctx.set_mu_tables(*_tables())
        pe, _ = ctx.pe_forward(0, "fp32")
        got, _ = ctx.scatter_crystal(pe, 0, "fp32")
        ref, _ = ctx.scatter_crystal(pe, 0, "fp64")
    _compare(got, ref, "crystal scatter, 2048 bars")


def test_collimator_grid_half_mm(smg_mod, params_mod):
    """The benchmark's collimator grid: 0.5 mm pitch (143,440 points), 32x32x16
    voxels x 64 bars, 64-node mu(E) tables (1,048,576 elements)."""
    p = params_mod.config(3)
    with smg_mod.open_context(p) as ctx:
        ctx.set_scatter_subsamples(*p.meta["subsamples"])
        ctx.set_mu_tables(*_tables())
        got, _, n = ctx.scatter_collimator_grid(0.5, 0, "fp32")
        ref, _, n64 = ctx.scatter_collimator_grid(0.5, 0, "fp64")
    assert n == n64 > 140000
    _compare(got, ref, "collimator grid 0.5 mm")


def test_crystal_grid_quarter_mm(smg_mod, params_mod):
    """The benchmark's crystal grid: 10x32x10 cells (0.25 mm) in each of the
    2048 bars (6,553,600 points), its 8 detector rows, 64-node mu(E) tables."""
    p = params_mod.config(4)
    rows = np.arange(0, 2048, 256, dtype=np.int32)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*p.meta["pe_subsamples"])
        ctx.set_scatter_subsamples(*p.meta["subsamples"])
        ctx.set_mu_tables(*_tables())
        got, _ = ctx.scatter_crystal_grid((10, 32, 10), rows, 0, "fp32")
        ref, _ = ctx.scatter_crystal_grid((10, 32, 10), rows, 0, "fp64")
    _compare(got, ref, "crystal grid 0.25 mm, 8 rows")
