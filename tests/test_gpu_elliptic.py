"""GPU parity with elliptical holes (smg_set_hole_shapes) against the FP64
oracle.  The chord is the reference's length_ellipticalcylinder_ray
(PEGen_RayTracing_CircularHole/PESysMatGen.cu:801-947), a routine the
reference defines but never calls; the oracle restates it and its pins are
in test_oracle_primitives.py."""
import numpy as np
import pytest

from conftest import TOL_FP32, TOL_FP64, assert_parity, elliptic_hole_shapes

pytestmark = pytest.mark.gpu


def _small(params_mod, n_det=512):
    """Full-size collimator (400 holes) and the first n_det bars, 8x8x2
    voxels at 2 mm: every element fits the oracle."""
    p = params_mod.config(2)
    p.det[0] = n_det
    params_mod.set_image(p, 8, 8, 2, 2.0, 2.0, 2.0)
    return p


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_pe_elliptic_holes_full_size(smg_mod, oracle_mod, params_mod, prec):
    """Full-size PE matrix with 300 of the 400 holes elliptical: every element
    on the GPU, a seeded sample of 20000 in the oracle.  An FP32 request runs
    the FP32 kernel with the elliptic chord compiled in (FP64 re-computation
    of its near rays, FP32 tolerance); FP64 the all-FP64 kernel."""
    p = params_mod.config(2)
    sub = p.meta["subsamples"]
    shapes = elliptic_hole_shapes(400, seed=5)
    o = oracle_mod.Oracle(p)
    o.set_hole_shapes(shapes)
    rng = np.random.default_rng(6)
    d = rng.integers(0, o.n_det, 20000)
    v = rng.integers(0, o.n_vox, 20000)
    ref = o.pe_pairs(d, v, sub=sub)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        circ, _ = ctx.pe_forward(0, prec)
        ctx.set_hole_shapes(shapes)
        got, st = ctx.pe_forward(0, prec)
    assert st["overflow_waves"] == 0
    if prec == "fp64":
        assert st["fp64_fallbacks"] == 0
    changed = np.abs(got[d, v] - circ[d, v]) > 1e-3 * np.abs(circ[d, v])
    assert changed.mean() > 0.3
    assert_parity(got[d, v], ref, TOL_FP64 if prec == "fp64" else TOL_FP32,
                  f"PE elliptic holes {prec} fallbacks={st['fp64_fallbacks']}")


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-7), ("fp64", 1e-7)])
def test_adjoint_elliptic_holes(smg_mod, oracle_mod, params_mod, prec, tol):
    """dL/dmu with elliptical holes (the holes' chords enter the plate path
    T/cos - sum of hole chords) against the FP64 host gradient."""
    p = _small(params_mod)
    sub = (3, 3, 3)
    shapes = elliptic_hole_shapes(400, seed=7)
    w = np.random.default_rng(8).uniform(0, 1, (p.n_det, p.n_vox)).astype(np.float32)
    o = oracle_mod.Oracle(p)
    o.set_hole_shapes(shapes)
    loss_ref, g_ref = o.pe_grad(w.astype(np.float64), sub=sub)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        ctx.set_hole_shapes(shapes)
        loss, g, _ = ctx.pe_backward(w, precision=prec)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref)
    print(f"adjoint elliptic holes {prec}: max |dg| / max |g| = {err.max() / scale:.2e}", flush=True)
    assert err.max() <= tol * scale
    assert abs(loss - loss_ref) <= tol * abs(loss_ref)


def test_projections_elliptic_holes(smg_mod, params_mod):
    """Matrix-free y = A x and x = A^T y equal the stored matrix's products."""
    p = _small(params_mod, 2048)
    shapes = elliptic_hole_shapes(400, seed=9)
    rng = np.random.default_rng(10)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(3, 3, 3)
        ctx.set_hole_shapes(shapes)
        A, _ = ctx.pe_forward(0, "fp32")
        x = rng.uniform(0, 1, ctx.dims.n_vox).astype(np.float32)
        y = rng.uniform(0, 1, ctx.dims.n_det).astype(np.float32)
        yp, _ = ctx.pe_project(x)
        xb, _ = ctx.pe_backproject(y)
    A64 = A.astype(np.float64)
    assert np.allclose(yp, A64 @ x, rtol=1e-6, atol=0)
    assert np.allclose(xb, A64.T @ y, rtol=1e-6, atol=0)


def test_hole_shapes_reset_and_errors(smg_mod, params_mod):
    """NULL shapes give back the circles bit for bit, a new geometry resets
    the shapes, and a wrong hole count or non-positive semi-axis is
    rejected."""
    p = _small(params_mod, 256)
    shapes = elliptic_hole_shapes(400, seed=11)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(2, 2, 2)
        circ, _ = ctx.pe_forward(0, "fp32")
        ctx.set_hole_shapes(shapes)
        ell, _ = ctx.pe_forward(0, "fp32")
        assert not np.array_equal(ell, circ)
        ctx.set_hole_shapes(None)
        again, _ = ctx.pe_forward(0, "fp32")
        assert np.array_equal(again, circ)
        ctx.set_hole_shapes(shapes)
        ctx.set_geometry(p)
        ctx.set_pe_subsamples(2, 2, 2)
        reset, _ = ctx.pe_forward(0, "fp32")
        assert np.array_equal(reset, circ)
        with pytest.raises(smg_mod.SmgError):
            ctx.set_hole_shapes(shapes[:399])
        bad = shapes.copy()
        bad[5, 1] = 0.0
        with pytest.raises(smg_mod.SmgError):
            ctx.set_hole_shapes(bad)
        # a failed call leaves the geometry as it was
        after, _ = ctx.pe_forward(0, "fp32")
        assert np.array_equal(after, circ)


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_collimator_grid_elliptic_holes(smg_mod, oracle_mod, params_mod, prec, tol):
    """Grid scatterers over the layer-0 tungsten (1 mm pitch, config-3
    geometry) with elliptical holes: the points inside the ellipses left out,
    the first and second legs through the plates less the elliptic-cylinder
    chords, against the oracle."""
    p = params_mod.config(3)
    n_h = p.n_holes_layer0
    rng = np.random.default_rng(24)
    shapes = np.zeros((n_h, 3), np.float32)
    shapes[:, 0] = rng.uniform(0.6, 1.4, n_h)
    shapes[:, 1] = rng.uniform(0.3, 0.7, n_h)
    shapes[:, 2] = rng.uniform(-np.pi, np.pi, n_h)
    shapes[::5, 0] = 0.0  # some circles
    o = oracle_mod.Oracle(p)
    o.set_hole_shapes(shapes)
    pts, _ = o.coll_grid_points(1.0)
    d = rng.integers(0, o.n_det, 200)
    v = rng.integers(0, o.n_vox, 200)
    with smg_mod.open_context(p) as ctx:
        circ, _, n_circ = ctx.scatter_collimator_grid(1.0, 0, prec)
        ctx.set_hole_shapes(shapes)
        got, st, n = ctx.scatter_collimator_grid(1.0, 0, prec)
    assert n == len(pts) and n != n_circ
    ref = o.coll_grid_scatter_pairs(1.0, d, v)
    assert (ref > 0).mean() > 0.5
    assert not np.allclose(got[d, v], circ[d, v], rtol=1e-3, atol=0.0)
    assert_parity(got[d, v], ref, tol, f"collimator grid, elliptical holes, {prec}")
