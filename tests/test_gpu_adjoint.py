"""GPU adjoint dL/dmu of L = sum w A (smg_pe_backward) against the FP64 host
gradient, central finite differences of the GPU forward, and a calibration
run.  The reference has no gradient; the derivatives follow from its
element formula (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1240-1357)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg5(params_mod):
    p = params_mod.config(5)
    w = np.random.default_rng(2).uniform(0, 1, (p.n_det, p.n_vox))
    return p, w


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("fp64", 1e-7)])
def test_gradient_matches_fp64_host(smg_mod, oracle_mod, cfg5, prec, tol):
    """FP32 path: components within 1e-5 of the largest and within 1e-4
    relative where above 1e-3 of it; all-FP64 path: 1e-7 and 1e-6."""
    p, w = cfg5
    sub = p.meta["subsamples"]
    loss_ref, g_ref = oracle_mod.Oracle(p).pe_grad(w, sub=sub)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        loss, g, st = ctx.pe_backward(w.astype(np.float32), precision=prec)
        A, _ = ctx.pe_forward(0, prec)
    w32 = w.astype(np.float32).astype(np.float64)
    assert loss == pytest.approx(float(np.sum(w32 * A.astype(np.float64))), rel=1e-6)
    _, g_ref32 = oracle_mod.Oracle(p).pe_grad(w32, sub=sub)
    scale = np.abs(g_ref32).max()
    err = np.abs(g - g_ref32)
    big = np.abs(g_ref32) > 1e-3 * scale
    print(f"adjoint {prec}: max |dg| / max |g| = {err.max() / scale:.2e}, max relative on the "
          f"large components {np.max(err[big] / np.abs(g_ref32[big])):.2e}, "
          f"kernel {st['kernel_ms']:.2f} ms", flush=True)
    assert err.max() <= tol * scale
    assert np.all(err[big] <= 10 * tol * np.abs(g_ref32[big]))
    assert np.abs(loss_ref - loss) <= tol * abs(loss_ref)


def test_gradient_matches_finite_differences(smg_mod, params_mod, cfg5):
    """Central differences of the all-FP64 GPU forward, h = 1e-4 relative,
    on 32 seeded parameters."""
    from smg import calibrate
    p, w = cfg5
    sub = p.meta["subsamples"]
    mu0 = calibrate.get_mu(p)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        _, g, _ = ctx.pe_backward(w.astype(np.float32))
    w32 = w.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(7)
    picks = np.concatenate([[0], rng.choice(np.arange(1, mu0.size), 31, replace=False)])

    def L(mu):
        with smg_mod.open_context(calibrate.set_mu(p, mu)) as ctx:
            ctx.set_pe_subsamples(*sub)
            A, _ = ctx.pe_forward(0, "fp64")
        return float(np.sum(w32 * A.astype(np.float64)))

    scale = np.abs(g).max()
    for i in picks:
        up, dn = mu0.copy(), mu0.copy()
        up[i] = np.float32(mu0[i] * (1 + 1e-4))
        dn[i] = np.float32(mu0[i] * (1 - 1e-4))
        fd = (L(up) - L(dn)) / (float(np.float32(up[i])) - float(np.float32(dn[i])))
        assert abs(fd - g[i]) <= 1e-4 * max(abs(g[i]), 1e-3 * scale), (i, fd, g[i])


def test_backward_deterministic(smg_mod, cfg5):
    p, w = cfg5
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(2, 2, 2)
        a = ctx.pe_backward(w.astype(np.float32))[1]
        b = ctx.pe_backward(w.astype(np.float32))[1]
    assert np.array_equal(a, b)


def _calibration_case(smg_mod, p, prec="fp32", sub=(2, 2, 2)):
    """1000 point sources (distinct voxels, seed 2), their projections at the
    true mu, and a start with every mu off by +-5%."""
    from smg import calibrate
    rng = np.random.default_rng(2)
    sources = np.sort(rng.choice(p.n_vox, 1000, replace=False))
    mu_true = calibrate.get_mu(p)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        A, _ = ctx.pe_forward(0, prec)
    target = A[:, sources].astype(np.float64)
    mu_start = mu_true * (1 + rng.choice([-0.05, 0.05], mu_true.size))
    return sub, sources, mu_true, target, calibrate.set_mu(p, mu_start)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_calibration_recovers_every_mu(smg_mod, params_mod, cfg5, prec):
    """Gauss-Newton with conjugate gradients on J^T (J p) products (tangent
    and adjoint passes, no per-parameter differences) from mu perturbed by
    +-5%, at config 5's 4x4x4 sub-cells: all 129 mu (plate, mu_total and
    mu_pe of 64 bars) recovered to 1e-4 within 20 steps."""
    from smg import calibrate
    p, _ = cfg5
    sub, sources, mu_true, target, p_start = _calibration_case(smg_mod, p, prec, sub=(4, 4, 4))
    res = calibrate.calibrate(p_start, target, sources, sub=sub, steps=20, precision=prec)
    err = np.abs(res.mu / mu_true - 1)
    print(f"calibration GN-CG {prec}: loss {res.losses[0]:.3e} -> {res.losses[-1]:.3e} in "
          f"{len(res.losses) - 1} steps ({res.n_forward} forward, {res.n_tangent} tangent, "
          f"{res.n_backward} adjoint passes, CG iterations {sum(res.cg_iterations)}), max mu error "
          f"0.050 -> {err.max():.2e} (plate {err[0]:.2e})", flush=True)
    assert all(b <= a for a, b in zip(res.losses, res.losses[1:]))
    assert err.max() < 1e-4
    assert res.n_forward + res.n_tangent + res.n_backward < 3000


def test_calibration_levenberg_marquardt_forward_differences(smg_mod, params_mod, cfg5):
    """The direct method for small parameter counts (J^T J from one forward
    difference per parameter) reaches the same mu at 2x2x2 sub-cells."""
    from smg import calibrate
    p, _ = cfg5
    sub, sources, mu_true, target, p_start = _calibration_case(smg_mod, p)
    res = calibrate.calibrate(p_start, target, sources, sub=sub, steps=20,
                              method="levenberg-marquardt")
    err = np.abs(res.mu / mu_true - 1)
    assert err.max() < 1e-3


def test_calibration_2048_bars(smg_mod, params_mod):
    """Gauss-Newton-CG on the full-size detector (2048 bars, 400-hole plate:
    4,097 mu) with 1000 point sources of a 16x16x8 image at 2x2x2 sub-cells:
    the passes per step do not grow with the parameter count (two per CG
    product), the loss falls by nine orders of magnitude and every mu is
    recovered."""
    from smg import calibrate
    p = params_mod.config(2)
    params_mod.set_image(p, 16, 16, 8, 0.5, 0.5, 0.5)
    sub, sources, mu_true, target, p_start = _calibration_case(smg_mod, p, "fp32", sub=(2, 2, 2))
    res = calibrate.calibrate(p_start, target, sources, sub=sub, steps=40)
    err = np.abs(res.mu / mu_true - 1)
    print(f"calibration 4097 mu: loss {res.losses[0]:.3e} -> {res.losses[-1]:.3e} in "
          f"{len(res.losses) - 1} steps ({res.n_forward} forward, {res.n_tangent} tangent, "
          f"{res.n_backward} adjoint passes), max mu error 0.050 -> {err.max():.2e}, median "
          f"{np.median(err):.2e}", flush=True)
    assert res.losses[-1] < 1e-9 * res.losses[0]
    assert err.max() < 1e-3


def test_calibration_adam_fits_projections(smg_mod, params_mod, cfg5):
    """200 Adam steps from the same start fit the projections and recover
    the plate mu; per-bar mu_total / mu_pe are weakly identifiable from these
    projections (Jacobian condition number ~1.6e4: the weakest direction
    trades a common scale of all mu_pe against the plate mu) and first-order
    steps stay within ~10% of them, so Adam's fit is checked in output
    space."""
    from smg import calibrate
    p, _ = cfg5
    sub, sources, mu_true, target, p_start = _calibration_case(smg_mod, p)
    res = calibrate.calibrate(p_start, target, sources, sub=sub, steps=200, method="adam")
    err = np.abs(res.mu / mu_true - 1)
    rms = np.sqrt(res.losses[-1] / target.size)
    print(f"calibration: loss {res.losses[0]:.3e} -> {res.losses[-1]:.3e} (rms rel residual "
          f"{rms:.2e}), plate mu error 0.050 -> {err[0]:.2e}, bar mu max error {err[1:].max():.3f}; "
          f"fwd {res.forward_ms:.0f} ms, bwd {res.backward_ms:.0f} ms", flush=True)
    assert res.losses[-1] < 1e-4 * res.losses[0]
    assert rms < 5e-3
    assert err[0] < 0.01


def test_full_size_adjoint_reduced_image(smg_mod, oracle_mod, params_mod):
    """The FP32 adjoint on the full-size geometry (2048 bars, 400 holes, 3x3x3
    sub-cells) with a 16x16x8 image at 0.5 mm, weights on the 24 rows of each
    detector layer with the largest PE row sums: against the FP64 host
    gradient, and central differences of the GPU FP64 forward on the plate mu
    and on the bars' mu_total and mu_pe with the largest gradients.  Some of
    its rays keep more than 8 narrowed bars (host lane check), so the chords
    the adjoint recomputes beyond the 8 it keeps in LDS are covered."""
    from smg import calibrate
    p = params_mod.config(2)
    params_mod.set_image(p, 16, 16, 8, 0.5, 0.5, 0.5)
    sub = p.meta["subsamples"]
    nb = p.n_det
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        A0, _ = ctx.pe_forward(0, "fp32")
    sums = A0.astype(np.float64).sum(axis=1)
    rows = np.sort(np.concatenate([np.argsort(sums[:1024])[-24:], 1024 + np.argsort(sums[1024:])[-24:]]))
    assert (sums[rows] > 0).all()
    rng = np.random.default_rng(62)
    w = np.zeros((nb, p.n_vox), np.float32)
    w[rows] = rng.uniform(0, 1, (len(rows), p.n_vox)).astype(np.float32)
    w64 = w.astype(np.float64)
    _, g_ref = oracle_mod.Oracle(p).pe_grad(w64, sub=sub)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        loss, g, st = ctx.pe_backward(w, precision="fp32")
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref)
    big = np.abs(g_ref) > 1e-3 * scale
    print(f"full-size adjoint, 16x16x8 image, {len(rows)} rows: max |dg| / max |g| = "
          f"{err.max() / scale:.2e}, max relative on the large components "
          f"{np.max(err[big] / np.abs(g_ref[big])):.2e}, kernel {st['kernel_ms']:.2f} ms", flush=True)
    assert err.max() <= 1e-5 * scale
    assert np.all(err[big] <= 1e-4 * np.abs(g_ref[big]))

    mu0 = calibrate.get_mu(p)

    def L(mu):
        with smg_mod.open_context(calibrate.set_mu(p, mu)) as c:
            c.set_pe_subsamples(*sub)
            A, _ = c.pe_forward(0, "fp64")
        return float(np.sum(w64 * A.astype(np.float64)))

    g_t, g_pe = np.abs(g[1:1 + nb]), np.abs(g[1 + nb:])
    picks = [0] + [1 + int(i) for i in np.argsort(g_t)[-3:]] + [1 + nb + int(i) for i in np.argsort(g_pe)[-2:]]
    for i in picks:
        up, dn = mu0.copy(), mu0.copy()
        up[i] = np.float32(mu0[i] * (1 + 1e-4))
        dn[i] = np.float32(mu0[i] * (1 - 1e-4))
        fd = (L(up) - L(dn)) / (float(np.float32(up[i])) - float(np.float32(dn[i])))
        print(f"  parameter {i}: FD {fd:.6e}, adjoint {g[i]:.6e}", flush=True)
        assert g[i] != 0.0
        assert abs(fd - g[i]) <= 1e-4 * abs(g[i]), (i, fd, g[i])


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("fp64", 1e-9)])
def test_tangent_is_the_adjoints_transpose(smg_mod, cfg5, prec, tol):
    """Dot-product test of the tangent (smg_pe_tangent, J v) against the
    adjoint (smg_pe_backward, J^T w): w . (J v) = (J^T w) . v for random w and
    v, and the tangent's matrix output is the forward's."""
    p, w = cfg5
    sub = p.meta["subsamples"]
    rng = np.random.default_rng(70)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        n_par = 1 + 2 * p.n_det
        for trial in range(3):
            v = rng.standard_normal(n_par) * np.abs(rng.uniform(0.1, 1.0, n_par))
            dA, A, st = ctx.pe_tangent(v, 0, prec, with_matrix=True)
            _, g, _ = ctx.pe_backward(w.astype(np.float32), precision=prec)
            lhs = float(np.sum(w.astype(np.float32).astype(np.float64) * dA.astype(np.float64)))
            rhs = float(g @ v)
            scale = float(np.sum(np.abs(w.astype(np.float32).astype(np.float64) * dA.astype(np.float64))))
            print(f"tangent {prec}: w.(J v) = {lhs:.9e}, (J^T w).v = {rhs:.9e}, "
                  f"kernel {st['kernel_ms']:.3f} ms", flush=True)
            assert abs(lhs - rhs) <= tol * scale
        A_ref, _ = ctx.pe_forward(0, prec)
    np.testing.assert_allclose(A, A_ref, rtol=1e-6, atol=0.0)


def test_tangent_matches_finite_differences(smg_mod, params_mod, cfg5):
    """The FP64 tangent along a direction over the plate mu and 8 bars' mu_total
    and mu_pe, element by element, against central differences of the FP64
    forward (h = 1e-3 relative on each coefficient: the forward's float32
    output leaves ~3e-5 of noise in the differences at that step)."""
    from smg import calibrate
    p, _ = cfg5
    sub = (2, 2, 2)
    mu0 = calibrate.get_mu(p)
    rng = np.random.default_rng(71)
    v = np.zeros(mu0.size)
    v[0] = mu0[0]
    for k in rng.choice(p.n_det, 8, replace=False):
        v[1 + k] = mu0[1 + k] * rng.uniform(-1, 1)
        v[1 + p.n_det + k] = mu0[1 + p.n_det + k] * rng.uniform(-1, 1)
    h = 1e-3

    def A_at(mu):
        with smg_mod.open_context(calibrate.set_mu(p, mu)) as c:
            c.set_pe_subsamples(*sub)
            a, _ = c.pe_forward(0, "fp64")
        return a.astype(np.float64)

    up, dn = mu0 + h * v, mu0 - h * v
    up32 = up.astype(np.float32).astype(np.float64)
    dn32 = dn.astype(np.float32).astype(np.float64)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*sub)
        dA, _, _ = ctx.pe_tangent((up32 - dn32) / (2 * h), 0, "fp64")
    fd = (A_at(up32) - A_at(dn32)) / (2 * h)
    scale = np.abs(fd).max()
    err = np.abs(dA.astype(np.float64) - fd)
    print(f"tangent vs FD: max |d| / max |FD| = {err.max() / scale:.2e}", flush=True)
    assert scale > 0 and err.max() <= 1e-4 * scale
