"""Matrix-free projections with the PE matrix (smg_pe_project,
smg_pe_backproject): y = A x and x = A^T y with each element formed as the
matrix kernel forms it and never stored.  The reference writes the matrix
for a reconstruction program to read (PEGen_RayTracing_CircularHole/
PEGen_CircularHole.cpp:141-147); the projections are its consumer."""
import os

import numpy as np
import pytest

from conftest import ROOT, TOL_FP32, TOL_FP64, assert_parity

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_projections_match_golden_matrix(smg_mod, params_mod, prec, tol):
    """Config 1 (32x32x1 voxels, 64 bars, 4x4x4 sub-cells) against the
    committed FP64 oracle matrix: y = A x and x = A^T y for positive x, y
    (sums of positive terms, so the element tolerance carries over)."""
    G = np.load(os.path.join(GOLDEN, "cfg1_pe_fp64.npy"))
    p = params_mod.config(1)
    rng = np.random.default_rng(31)
    x = rng.uniform(0.5, 2.0, p.n_vox).astype(np.float32)
    y = rng.uniform(0.5, 2.0, p.n_det).astype(np.float32)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(4, 4, 4)
        yp, st = ctx.pe_project(x, 0, prec)
        xb, _ = ctx.pe_backproject(y, 0, prec)
    assert_parity(yp, G @ x.astype(np.float64), tol, f"forward projection {prec}")
    assert_parity(xb, G.T @ y.astype(np.float64), tol, f"back-projection {prec}")


@pytest.mark.parametrize("cfg,sub,rot", [(1, (4, 4, 4), 0), (1, (2, 3, 2), 0), ("rot", (2, 2, 2), 1)])
def test_projections_equal_stored_matrix(smg_mod, params_mod, cfg, sub, rot):
    """The projections equal the stored matrix (smg_pe_forward) times the
    vector, up to the FP64 summation order; zeros in y skip bars; repeated
    calls give the same bits."""
    if cfg == "rot":
        p = params_mod.config(1)
        params_mod.set_image(p, 16, 16, 4, 1.0, 1.0, 1.0, n_rot=2, dphi=0.3)
    else:
        p = params_mod.config(cfg)
    rng = np.random.default_rng(32)
    x = rng.uniform(-1.0, 2.0, p.n_vox).astype(np.float32)
    y = rng.uniform(0.0, 2.0, p.n_det).astype(np.float32)
    y[rng.choice(p.n_det, p.n_det // 3, replace=False)] = 0.0
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        A, _ = ctx.pe_forward(rot, "fp32")
        yp, _ = ctx.pe_project(x, rot, "fp32")
        yp2, _ = ctx.pe_project(x, rot, "fp32")
        xb, _ = ctx.pe_backproject(y, rot, "fp32")
    A = A.astype(np.float64)
    np.testing.assert_allclose(yp, A @ x.astype(np.float64), rtol=2e-7, atol=1e-7 * np.abs(A).sum(1).max())
    np.testing.assert_allclose(xb, A.T @ y.astype(np.float64), rtol=2e-7, atol=0)
    assert np.array_equal(yp, yp2)


def test_projections_full_size(smg_mod, params_mod):
    """The benchmark geometry (64x64x32 voxels, 2048 bars, 400 holes, 3x3x3
    sub-cells): projections against the stored 1 GiB matrix, and the adjoint
    identity <A x, y> = <x, A^T y>."""
    p = params_mod.config(2)
    rng = np.random.default_rng(33)
    x = rng.uniform(0.0, 1.0, p.n_vox).astype(np.float32)
    y = rng.uniform(0.0, 1.0, p.n_det).astype(np.float32)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(3, 3, 3)
        yp, st_f = ctx.pe_project(x, 0, "fp32")
        xb, st_b = ctx.pe_backproject(y, 0, "fp32")
        A, st_m = ctx.pe_forward(0, "fp32")
    print(f"full size: matrix {st_m['kernel_ms']:.1f} ms, project {st_f['kernel_ms']:.1f} ms, "
          f"back-project {st_b['kernel_ms']:.1f} ms", flush=True)
    ref_y = A.astype(np.float64) @ x.astype(np.float64)
    np.testing.assert_allclose(yp, ref_y, rtol=1e-6, atol=0)
    ref_x = y.astype(np.float64) @ A.astype(np.float64)
    np.testing.assert_allclose(xb, ref_x, rtol=1e-6, atol=1e-12 * ref_x.max())
    lhs = float(yp.astype(np.float64) @ y.astype(np.float64))
    rhs = float(x.astype(np.float64) @ xb.astype(np.float64))
    assert lhs == pytest.approx(rhs, rel=1e-6)
