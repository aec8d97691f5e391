"""Element-level pins of the FP64 oracle: properties the reference formula
must satisfy whatever its implementation (PEGen_RayTracing_CircularHole/
PESysMatGen.cu:951-1364; ScatterGen_RayTracing_CircularHole/scatter.cu:1077-1936)."""
import math

import numpy as np
import pytest


def one_bar_geometry(params_mod, mu_coll=3.65, plate_w=200.0, hole=None, bar=(2.0, 10.0, 2.0),
                     bar_mu=(0.25, 0.18, 0.07), y_bar=17.5, F=30.0, n_img=(1, 1, 1), w_img=1.0):
    p = params_mod.Params.empty()
    p.coll[0] = 1
    params_mod.set_layer(p, 0, 1 if hole else 0, plate_w, 5.0, plate_w, 0.0, mu_coll, 0.9 * mu_coll,
                         0.1 * mu_coll)
    if hole:
        params_mod.set_hole(p, 0, hole[0], -2.5, 2.5, hole[1], hole[2])
    p.det[0] = 1
    params_mod.set_bar(p, 0, 0.0, y_bar, 0.0, *bar, *bar_mu, 0.1)
    params_mod.set_image(p, *n_img, w_img, w_img, w_img, fov2coll=F)
    params_mod.set_physics(p)
    return p


def test_thin_bar_limit(oracle_mod, params_mod):
    """No collimator material, a bar far away on the voxel's axis: the
    sub-cell sum is the absorption integral through a slab,
    w h (1 - exp(-mu t)) / (4 pi D^2) * mu_pe / mu_total."""
    mu = 0.02
    p = one_bar_geometry(params_mod, mu_coll=0.0, bar_mu=(mu, mu, 0.0), y_bar=400.0)
    o = oracle_mod.Oracle(p)
    A = o.pe_element(0, 0, sub=(4, 8, 4))
    D = 430.0  # voxel at the origin, bar centre at y = F + 400
    expect = 2.0 * 2.0 * (1 - math.exp(-mu * 10.0)) / (4 * math.pi * D * D)
    assert A == pytest.approx(expect, rel=3e-3)


def test_plate_attenuation_scaling(oracle_mod, params_mod):
    """Without holes every ray crosses T / cos(alpha) of plate: for a voxel on
    the bar axis d ln A / d mu_coll = -T / cos(alpha) ~ -T."""
    mus = (3.0, 3.1)
    vals = []
    for m in mus:
        p = one_bar_geometry(params_mod, mu_coll=m)
        vals.append(oracle_mod.Oracle(p).pe_element(0, 0, sub=(2, 4, 2)))
    slope = (math.log(vals[1]) - math.log(vals[0])) / (mus[1] - mus[0])
    assert slope == pytest.approx(-5.0, rel=2e-3)


def test_blocked_when_plate_not_crossed(oracle_mod, params_mod):
    """A ray that does not cross the plate box gets +1e5 and contributes
    exactly 0 (PESysMatGen.cu:1251-1253): a plate 1 mm wide misses the rays
    from a voxel 10 mm to the side."""
    p = one_bar_geometry(params_mod, plate_w=1.0)
    p.img[8] = 10.0  # shift the voxel in x
    assert oracle_mod.Oracle(p).pe_element(0, 0, sub=(2, 2, 2)) == 0.0


def test_hole_on_axis_ignores_plate_mu(oracle_mod, params_mod):
    """A voxel on the axis of a wide hole sees the bar only through the hole:
    the plate's mu does not matter."""
    vals = []
    for m in (3.65, 365.0):
        p = one_bar_geometry(params_mod, mu_coll=m, hole=(0.0, 0.0, 3.0))
        vals.append(oracle_mod.Oracle(p).pe_element(0, 0, sub=(2, 4, 2)))
    assert vals[0] > 0 and vals[0] == pytest.approx(vals[1], rel=1e-14)


def test_mirror_symmetry(oracle_mod, params_mod):
    """A geometry symmetric under x -> -x gives A[d][v] = A[d'][v'] for the
    mirrored bar and voxel."""
    p = params_mod.Params.empty()
    p.coll[0] = 1
    params_mod.set_layer(p, 0, 4, 40.0, 5.0, 40.0, 0.0, 3.65, 3.3, 0.35)
    for h, (x, z) in enumerate([(-3.0, 1.0), (3.0, 1.0), (-1.5, -2.0), (1.5, -2.0)]):
        params_mod.set_hole(p, h, x, -2.5, 2.5, z, 0.7)
    xs = [-3.0, -1.0, 1.0, 3.0]
    p.det[0] = 8
    for k in range(8):
        params_mod.set_bar(p, k, xs[k % 4], 17.5, -1.0 + 2.0 * (k // 4), 2.0, 10.0, 2.0,
                           0.25, 0.18, 0.07, 0.1)
    params_mod.set_image(p, 6, 5, 2, 1.0, 1.0, 1.0)
    params_mod.set_physics(p)
    o = oracle_mod.Oracle(p)
    A = o.pe_matrix(sub=(2, 3, 2)).reshape(8, 2, 5, 6)
    mirror_bar = [3, 2, 1, 0, 7, 6, 5, 4]
    assert np.all(A > 0)
    np.testing.assert_allclose(A, A[mirror_bar][..., ::-1], rtol=1e-12)


def test_subcell_additivity(oracle_mod, params_mod):
    """Splitting every sub-cell along y into two halves: the element is a sum
    over sub-cells of smooth terms, so 2 x 8 x 2 and 2 x 16 x 2 agree closely."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    a = o.pe_element(27, 500, sub=(2, 8, 2))
    b = o.pe_element(27, 500, sub=(2, 16, 2))
    assert a == pytest.approx(b, rel=2e-2)


def test_culling_counts_cfg2(oracle_mod, params_mod):
    """Culling keeps a few dozen bars and about ten holes per element of the
    full-size geometry (the reference's O(N + H) scan reduces to these)."""
    o = oracle_mod.Oracle(params_mod.config(2))
    nc, nh = o.cull_counts(1500, 60007)
    assert 20 < nc < 200 and 2 < nh < 40


# ---- scatter -------------------------------------------------------------------------

def test_crystal_scatter_linear_in_pe(oracle_mod, params_mod):
    """The crystal scatter is linear in the PE block (scatter.cu:1218-1219)."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    bm = o.bitmap_crystal()
    rng = np.random.default_rng(1)
    pe1 = rng.uniform(0, 1e-9, (64, 1024)).astype(np.float32)
    pe2 = rng.uniform(0, 1e-9, (64, 1024)).astype(np.float32)
    d = rng.integers(0, 64, 60)
    v = rng.integers(0, 1024, 60)
    s1 = o.crystal_scatter_pairs(pe1, bm, d, v)
    s2 = o.crystal_scatter_pairs(pe2, bm, d, v)
    s12 = o.crystal_scatter_pairs((pe1.astype(np.float64) + pe2).astype(np.float32), bm, d, v)
    np.testing.assert_allclose(s12, s1 + s2, rtol=2e-7)
    assert (s1 > 0).any()


def test_collimator_scatter_scales_with_compton_fraction(oracle_mod, params_mod):
    """The hole weight is Omega * mu_c / mu_total of layer 0 (scatter.cu:1716)."""
    p = params_mod.config(1)
    o1 = oracle_mod.Oracle(p)
    bm = o1.bitmap_collimator()
    q = p.copy()
    q.coll[17] *= 2.0
    o2 = oracle_mod.Oracle(q)
    d = np.arange(0, 64, 7)
    v = np.arange(0, 1024, 113)[: d.size]
    a = o1.collimator_scatter_pairs(bm, d, v)
    b = o2.collimator_scatter_pairs(bm, d, v)
    assert (a > 0).all()
    np.testing.assert_allclose(b, 2 * a, rtol=1e-14)


def test_collimator_variants_ratio(oracle_mod, params_mod):
    """Combined (r = 0.1, pi and cos factor) and standalone (r = 0.2) weights
    differ by a known per-(voxel, hole) factor; with a single hole the
    element ratio is that factor (scatter.cu:1711-1716 vs
    ScatterGen_Collimator/scatter.cu:1212-1215)."""
    p = params_mod.config(1)
    p.coll[10] = 1  # one hole
    o = oracle_mod.Oracle(p)
    bm = o.bitmap_collimator()
    d, v = np.array([20, 35]), np.array([400, 700])
    a = o.collimator_scatter_pairs(bm, d, v, variant=0)
    b = o.collimator_scatter_pairs(bm, d, v, variant=1)
    hx, hz = float(p.coll[100]), float(p.coll[103])
    for i in range(2):
        iy, ix = divmod(int(v[i]), 32)
        pv = np.array([ix - 15.5, iy - 15.5, 0.0])
        qh = np.array([hx, 30.0, hz])
        L = np.linalg.norm(qh - pv)
        r1, r2, T = float(np.float32(0.1)), float(np.float32(0.2)), 5.0
        w1 = (r1 * r1 * math.pi + T * r1) / (4 * math.pi * L * L) * abs(qh[1] - pv[1]) / L
        w2 = (r2 * r2 + T * r2) / (4 * L * L)
        if a[i] > 0:
            assert a[i] / b[i] == pytest.approx(w1 / w2, rel=1e-12)


def test_window_excludes_far_energies(oracle_mod, params_mod):
    """A window far above E0 keeps no scattered photon: every term is cut by
    the 2-sigma early out (scatter.cu:1263-1269)."""
    p = params_mod.config(1)
    p.phys[5], p.phys[6] = 300.0, 400.0
    o = oracle_mod.Oracle(p)
    bm = o.bitmap_collimator()
    out = o.collimator_scatter_pairs(bm, np.arange(64), np.arange(64) * 16)
    assert np.all(out == 0.0)


def test_scatter_op_counter_counts_terms(oracle_mod, params_mod):
    """The scatter op counter visits every (scatter point, detection cell)
    term of the sampled elements once: the 63 other bar centres, the 16 hole
    centres, the grid cells of the other bars, times the detection cells."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    d = np.array([3, 40], np.int32)
    v = np.array([100, 9000], np.int32)
    bm_c = o.bitmap_crystal()
    bm_l = o.bitmap_collimator()
    ops0, t0 = o.scatter_ops_per_term(0, d, v, bitmap=bm_c)
    ops1, t1 = o.scatter_ops_per_term(1, d, v, bitmap=bm_l)
    ops3, t3 = o.scatter_ops_per_term(3, d, v, bitmap=bm_c, grid=(1, 2, 1))
    assert t0 == 2 * 63 and t1 == 2 * 16 and t3 == 2 * 63 * 2
    ops_s, t_s = o.scatter_ops_per_term(0, d, v, bitmap=bm_c, sub=(2, 1, 2))
    assert t_s == 4 * t0
    # the cheapest term is the window-rejected one (cell, cone angle, E', tests)
    assert 60 < min(ops0, ops1, ops3)
    assert ops0 < 20000 and ops1 < 20000


def test_pe_op_counter_excludes_culling_scan(oracle_mod, params_mod):
    """The PE op count is per sub-cell ray and leaves out the element's
    culling scan: bars far behind the detector, which no element culls in,
    add scan work per element but no operation to any ray."""
    p = params_mod.config(1)
    q = params_mod.config(1)
    n = p.n_det
    extra = 200
    q.det[0] = n + extra
    for i in range(extra):
        r = 12 * (n + i)
        q.det[r + 1:r + 13] = q.det[1:13]
        q.det[r + 1] = 5000.0 + 3.0 * i  # x far off to the side
        q.det[r + 2] = 9000.0            # y far behind
    d = np.arange(0, n, 9, dtype=np.int32)
    v = (np.arange(d.size) * 97) % p.n_vox
    sub = (2, 2, 2)
    a = oracle_mod.Oracle(p).pe_ops_per_ray(d, v, sub=sub)
    b = oracle_mod.Oracle(q).pe_ops_per_ray(d, v, sub=sub)
    assert a == b and 100 < a < 10000
