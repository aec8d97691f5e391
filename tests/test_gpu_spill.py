"""Tiles whose candidate lists exceed the PE kernels' LDS lists (448 bars, 320 in the adjoint,
256 holes; the reference culls up to 3000 bars and 1000 holes per element,
PEGen_RayTracing_CircularHole/PESysMatGen.cu:1087,1094): the kernels queue
them and redo them with lists in global memory.  Every PE-family entry point
is checked on such a geometry against the FP64 oracle, with the spilled work
items counted in the stats."""
import numpy as np
import pytest

from conftest import TOL_FP32, TOL_FP64, assert_parity, overflow_geometry

pytestmark = pytest.mark.gpu

SUB = (1, 2, 1)


@pytest.fixture(scope="module")
def geo(params_mod):
    return overflow_geometry(params_mod)


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_spilled_matrix(smg_mod, oracle_mod, geo, prec, tol):
    p = geo
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(45)
    # half the sample in the back layers, whose tiles overflow most
    d = np.concatenate([rng.integers(0, o.n_det, 1500), rng.integers(o.n_det * 3 // 4, o.n_det, 1500)])
    v = rng.integers(0, o.n_vox, 3000)
    ref = o.pe_pairs(d, v, sub=SUB)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*SUB)
        got, st = ctx.pe_forward(0, prec)
        again, _ = ctx.pe_forward(0, prec)
    print(f"overflow geometry {prec}: {st['overflow_waves']} of "
          f"{p.n_det * 8} waves redone with spill lists", flush=True)
    assert st["overflow_waves"] > 0
    assert np.isfinite(got).all()
    assert np.array_equal(got, again)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[d, v], ref, tol, f"spilled PE {prec}")


def test_spilled_projections(smg_mod, geo):
    """Forward and back-projection with spilled tiles equal the stored matrix
    times the vector."""
    p = geo
    rng = np.random.default_rng(46)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*SUB)
        A, _ = ctx.pe_forward(0, "fp32")
        x = rng.uniform(0, 1, p.n_vox).astype(np.float32)
        y = rng.uniform(0, 1, p.n_det).astype(np.float32)
        fy, st_f = ctx.pe_project(x, 0, "fp32")
        bx, st_b = ctx.pe_backproject(y, 0, "fp32")
    assert st_f["overflow_waves"] > 0 and st_b["overflow_waves"] > 0
    A64 = A.astype(np.float64)
    assert_parity(fy, A64 @ x.astype(np.float64), 1e-6, "spilled forward projection")
    assert_parity(bx, A64.T @ y.astype(np.float64), 1e-6, "spilled back-projection")


def test_spilled_adjoint(smg_mod, oracle_mod, geo):
    """The adjoint with spilled work items against the FP64 host gradient."""
    p = geo
    sub = (1, 1, 1)
    w = np.random.default_rng(47).uniform(0, 1, (p.n_det, p.n_vox))
    w32 = w.astype(np.float32)
    _, g_ref = oracle_mod.Oracle(p).pe_grad(w32.astype(np.float64), sub=sub)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        loss, g, st = ctx.pe_backward(w32, precision="fp32")
    assert st["overflow_waves"] > 0
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref)
    print(f"spilled adjoint: max |dg| / max |g| = {err.max() / scale:.2e}", flush=True)
    assert err.max() <= 1e-5 * scale


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_spilled_crystal_grid(smg_mod, oracle_mod, params_mod, prec, tol):
    """Crystal grid scatterers, whose first-leg weights come from the PE
    kernel's point mode, on a 4-layer version of the geometry (its tiles
    overflow the hole lists)."""
    p = overflow_geometry(params_mod, n_layers=4)
    o = oracle_mod.Oracle(p)
    grid = (1, 2, 1)
    rows = np.array([500, 560], np.int32)
    rng = np.random.default_rng(48)
    ri = rng.integers(0, len(rows), 200)
    v = rng.integers(0, o.n_vox, 200)
    with smg_mod.open_context(p) as ctx:
        got, st = ctx.scatter_crystal_grid(grid, rows, 0, prec)
        bm = ctx.bitmap_crystal()
    assert st["overflow_waves"] > 0
    ref = o.crystal_grid_scatter_pairs(bm, grid, rows[ri], v)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[ri, v], ref, tol, f"spilled crystal grid {prec}")
