"""Pins the FP64 oracle's geometry and physics primitives against closed
forms, brute force and library routines (the reference ships no tests or
golden data, so these are what fixes the oracle).

References restated by the oracle:
  length_box_ray          PEGen_RayTracing_CircularHole/PESysMatGen.cu:20-312
  length_box_ray_inside   PESysMatGen.cu:314-686
  length_cylinder_ray     PESysMatGen.cu:688-799
  length_ellipticalcylinder_ray   PESysMatGen.cu:801-947 (never called by the reference)
  KN / KN sum / cone angle / E' / window   ScatterGen_RayTracing_CircularHole/scatter.cu:961-1038
"""
import math

import numpy as np
import pytest
from scipy import special

EPS = float(np.float32(0.001))


def clip_chord(a, b, lo, hi):
    """Independent chord of segment a->b in a box: Liang-Barsky clipping of
    the parameter s in [0, 1], returning (chord, s_in, s_out) or None."""
    a, b, lo, hi = (np.asarray(x, float) for x in (a, b, lo, hi))
    d = b - a
    s0, s1 = -np.inf, np.inf
    for k in range(3):
        if d[k] == 0.0:
            if a[k] < lo[k] or a[k] > hi[k]:
                return None
            continue
        t1, t2 = (lo[k] - a[k]) / d[k], (hi[k] - a[k]) / d[k]
        s0, s1 = max(s0, min(t1, t2)), min(s1, max(t1, t2))
    if s0 > s1:
        return None
    L = np.linalg.norm(d)
    return (s1 - s0) * L, s0, s1


def test_box_ray_generic_random(oracle_mod):
    """Generic rays that fully cross a box: the chord equals the clipped
    parameter interval times the segment length."""
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(3000):
        lo = rng.uniform(-5, 0, 3)
        hi = lo + rng.uniform(0.5, 5, 3)
        c = rng.uniform(lo - 0.5, hi + 0.5)  # aim through (or near) the box
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = c - u * rng.uniform(1, 30)
        b = c + u * rng.uniform(1, 30)
        if np.min(np.abs(b - a)) < 0.01:
            continue
        r = clip_chord(a, b, lo, hi)
        got = oracle_mod.box_ray(a, b, lo, hi)
        if r is None or r[1] <= 0.0 or r[2] >= 1.0:  # missed, or an end point inside
            assert got == 0.0
            continue
        chord = r[0]
        L = np.linalg.norm(b - a)
        if chord < EPS or r[1] * L <= EPS:
            continue  # threshold cases are pinned separately below
        assert got == pytest.approx(chord, rel=1e-12, abs=1e-12)
        checked += 1
    assert checked > 300


def test_box_ray_axis_rays_exact_extent(oracle_mod):
    """A segment along one axis (other components below eps) through the
    whole box gives exactly the box extent (PESysMatGen.cu:35-66)."""
    lo, hi = [-1.0, -2.0, -0.5], [1.0, 3.0, 0.5]
    assert oracle_mod.box_ray([0.2, -9, 0.1], [0.2, 9, 0.1], lo, hi) == 5.0
    assert oracle_mod.box_ray([-9, 0.0, 0.2], [9, 0.0005, 0.2], lo, hi) == 2.0
    assert oracle_mod.box_ray([0.0, 0.0, 7], [0.0, 0.0, -7], lo, hi) == 1.0
    # the segment must run past both faces
    assert oracle_mod.box_ray([0.2, -9, 0.1], [0.2, 1.0, 0.1], lo, hi) == 0.0
    # outside the cross-section
    assert oracle_mod.box_ray([1.5, -9, 0.1], [1.5, 9, 0.1], lo, hi) == 0.0


def test_box_ray_one_flat_component(oracle_mod):
    """One component below eps: the 2-D chord, and 0 unless the start lies
    strictly between that axis's faces (PESysMatGen.cu:69-123)."""
    lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    a, b = [0.3, -5.0, -5.0], [0.3, 5.0, 5.0]
    assert oracle_mod.box_ray(a, b, lo, hi) == pytest.approx(2.0 * math.sqrt(2.0), rel=1e-14)
    assert oracle_mod.box_ray([1.0, -5, -5], [1.0, 5, 5], lo, hi) == 0.0  # on the face
    # the tiny x component is ignored: same chord as the flat ray
    assert oracle_mod.box_ray([0.3, -5, -5], [0.3005, 5, 5], lo, hi) == pytest.approx(
        2.0 * math.sqrt(2.0), rel=1e-14)


def test_box_ray_segment_ends_inside_is_zero(oracle_mod):
    """Both slab parameters must lie in (eps, |b-a|): starting or ending inside
    the box gives 0 (PESysMatGen.cu:297-304)."""
    lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    assert oracle_mod.box_ray([0.1, 0.2, 0.3], [5, 7, 9], lo, hi) == 0.0
    assert oracle_mod.box_ray([-5, -7, -9], [0.1, 0.2, 0.3], lo, hi) == 0.0
    assert oracle_mod.box_ray([-5, -7, -9], [5, 7, 9], lo, hi) > 0.0


def test_box_ray_short_chord_dropped(oracle_mod):
    """A chord below eps = 0.001 counts as 0 (PESysMatGen.cu:297-298)."""
    lo, hi = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
    # clip the corner: x + y crossing near (1, 1)
    a = [-10.0, 1.0 - 0.0004 - 10.0, 0.5]
    b = [10.0, 1.0 - 0.0004 + 10.0, 0.6]
    r = clip_chord(a, b, lo, hi)
    assert r is not None and r[0] < EPS
    assert oracle_mod.box_ray(a, b, lo, hi) == 0.0


def test_box_ray_inside_clips_at_end_point(oracle_mod):
    """Segment ending inside the box: chord from the entry face to the end
    point (PESysMatGen.cu:415-434)."""
    rng = np.random.default_rng(2)
    lo, hi = np.array([-1.0, -5.0, -1.0]), np.array([1.0, 5.0, 1.0])
    for _ in range(500):
        b = rng.uniform(lo * 0.9, hi * 0.9)
        a = b + rng.uniform(-30, 30, 3)
        a[1] = b[1] - rng.uniform(20, 40)
        if np.min(np.abs(b - a)) < 0.01:
            continue
        r = clip_chord(a, b, lo, hi)
        assert r is not None and r[2] == pytest.approx(1.0) or r[2] > 1.0
        expect = (1.0 - r[1]) * np.linalg.norm(b - a)
        assert oracle_mod.box_ray_inside(a, b, lo, hi) == pytest.approx(expect, rel=1e-11)
    # |dy| < 0.001 gives 0 (PESysMatGen.cu:321-324)
    assert oracle_mod.box_ray_inside([-5, 0.0, 0.2], [0.5, 0.0005, 0.2], lo, hi) == 0.0
    # along y with x, z inside: end point minus the entry face
    assert oracle_mod.box_ray_inside([0.2, -20, 0.1], [0.2, 1.5, 0.1], lo, hi) == 6.5


def cyl_chord_brute(a, b, xc, y1, y2, zc, r, n=200001):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.linspace(0.0, 1.0, n)
    pts = a[None, :] + s[:, None] * (b - a)[None, :]
    inside = ((pts[:, 0] - xc) ** 2 + (pts[:, 2] - zc) ** 2 <= r * r) & (pts[:, 1] >= y1) & (pts[:, 1] <= y2)
    return inside.mean() * np.linalg.norm(b - a), np.linalg.norm(b - a) / (n - 1)


def test_cylinder_through_both_caps(oracle_mod):
    """A ray crossing both end discs: chord = (y2 - y1)/|k_y|."""
    a, b = np.array([0.1, -20.0, -0.2]), np.array([0.5, 40.0, 0.3])
    k = (b - a) / np.linalg.norm(b - a)
    got = oracle_mod.cylinder_ray(a, b, 0.3, 10.0, 15.0, 0.05, 0.5)
    assert got == pytest.approx(5.0 / k[1], rel=1e-13)


def test_cylinder_misses_and_tangent(oracle_mod):
    a, b = [3.0, -20.0, 0.0], [3.0, 40.0, 0.0]
    assert oracle_mod.cylinder_ray(a, b, 0.0, 10.0, 15.0, 0.0, 0.5) == 0.0
    # just inside tangency: discriminant below 1e-5 gives 0 (PESysMatGen.cu:745)
    a, b = [0.5 - 1e-7, -20.0, -5.0], [0.5 - 1e-7, 40.0, 5.0]
    assert oracle_mod.cylinder_ray(a, b, 0.0, 10.0, 15.0, 0.0, 0.5) == 0.0


def test_cylinder_wall_crossings_brute_force(oracle_mod):
    """Oblique rays entering or leaving through the wall, against dense
    sampling of the segment."""
    rng = np.random.default_rng(3)
    n_wall = 0
    for _ in range(150):
        xc, zc, r = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.3, 1.0)
        a = np.array([xc + rng.uniform(-3, 3), -10.0, zc + rng.uniform(-3, 3)])
        b = np.array([xc + rng.uniform(-3, 3), 30.0, zc + rng.uniform(-3, 3)])
        got = oracle_mod.cylinder_ray(a, b, xc, 8.0, 14.0, zc, r)
        ref, step = cyl_chord_brute(a, b, xc, 8.0, 14.0, zc, r)
        assert got == pytest.approx(ref, abs=3 * step + 1e-9)
        n_wall += 0.0 < got < 6.0 * np.linalg.norm(b - a) / 40.0 - 1e-6
    assert n_wall > 5


def ell_chord_brute(a, b, xc, y1, y2, zc, xr, zr, theta, n=200001):
    """Length of the segment inside the turned elliptic cylinder, by sampling."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.linspace(0.0, 1.0, n)
    pts = a[None, :] + s[:, None] * (b - a)[None, :]
    dx, dz = pts[:, 0] - xc, pts[:, 2] - zc
    xp = dx * math.cos(theta) + dz * math.sin(theta)
    zp = -dx * math.sin(theta) + dz * math.cos(theta)
    inside = ((xp / xr) ** 2 + (zp / zr) ** 2 <= 1.0) & (pts[:, 1] >= y1) & (pts[:, 1] <= y2)
    return inside.mean() * np.linalg.norm(b - a), np.linalg.norm(b - a) / (n - 1)


def test_ellcyl_brute_force(oracle_mod):
    """Elliptic-cylinder chord (PESysMatGen.cu:801-947) of segments whose ends
    lie beyond both caps, against dense sampling of the segment: wall
    crossings, cap crossings and misses, turned ellipses of aspect up to 3.
    Semi-axes and obliquity are such that the discriminant cut (next test)
    only acts next to tangency."""
    rng = np.random.default_rng(11)
    kinds = {"wall": 0, "caps": 0, "miss": 0}
    for _ in range(200):
        xc, zc = rng.uniform(-1, 1), rng.uniform(-1, 1)
        xr, zr = rng.uniform(0.8, 2.5), rng.uniform(0.8, 2.5)
        th = rng.uniform(-math.pi, math.pi)
        a = np.array([xc + rng.uniform(-5, 5), -10.0, zc + rng.uniform(-5, 5)])
        b = np.array([xc + rng.uniform(-5, 5), 30.0, zc + rng.uniform(-5, 5)])
        got = oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, xr, zr, th)
        ref, step = ell_chord_brute(a, b, xc, 8.0, 14.0, zc, xr, zr, th)
        assert got == pytest.approx(ref, abs=3 * step + 1e-9), (xc, zc, xr, zr, th)
        full = 6.0 * np.linalg.norm(b - a) / 40.0
        kinds["miss" if got == 0.0 else ("caps" if abs(got - full) < 1e-9 else "wall")] += 1
    assert min(kinds.values()) > 5, kinds


def test_ellcyl_discriminant_cut_in_ellipse_units(oracle_mod):
    """The routine cuts b^2 - a c <= 1e-5 with its quadratic a = kx^2 z_rho^2 +
    kz^2 x_rho^2, ... (PESysMatGen.cu:887-893), which is x_rho^2 z_rho^2 times
    the unit-circle form: a ray that crosses the wall of a small ellipse almost
    along its axis gets 0.  The decision follows the discriminant formed
    independently here (in the turned frame, Lagrange form)."""
    rng = np.random.default_rng(13)
    n_cut = n_kept = 0
    for i in range(600):
        xc, zc = rng.uniform(-1, 1), rng.uniform(-1, 1)
        xr, zr, th = rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(-3, 3)
        w = 1.0 if i % 2 else 4.0  # near-axis and oblique rays
        a = np.array([xc + rng.uniform(-w, w), -10.0, zc + rng.uniform(-w, w)])
        b = np.array([xc + rng.uniform(-w, w), 30.0, zc + rng.uniform(-w, w)])
        got = oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, xr, zr, th)
        ref, step = ell_chord_brute(a, b, xc, 8.0, 14.0, zc, xr, zr, th)
        full = 6.0 * np.linalg.norm(b - a) / 40.0
        if ref == 0.0 or abs(ref - full) < 3 * step:
            continue  # misses and cap-to-cap rays do not reach the quadratic
        c, s = math.cos(th), math.sin(th)
        d = (b - a) / np.linalg.norm(b - a)
        kx, kz = d[0] * c + d[2] * s, -d[0] * s + d[2] * c
        x0 = (a[0] - xc) * c + (a[2] - zc) * s
        z0 = -(a[0] - xc) * s + (a[2] - zc) * c
        disc = xr * xr * zr * zr * (kx * kx * zr * zr + kz * kz * xr * xr - (kx * z0 - kz * x0) ** 2)
        if abs(disc - 1e-5) < 1e-9:
            continue
        if disc <= 1e-5:
            assert got == 0.0
            n_cut += 1
        else:
            assert got == pytest.approx(ref, abs=3 * step + 1e-9)
            n_kept += 1
    assert n_cut > 5 and n_kept > 5, (n_cut, n_kept)


def test_ellcyl_circle_and_axis_swap(oracle_mod):
    """x_rho = z_rho = 1, theta = 0 is the circular chord, bit for bit up to
    rounding, including rays next to tangency (the discriminant cut of both is
    1e-5 when r = 1); any theta turns a circle onto itself; (x_rho, z_rho,
    theta) and (z_rho, x_rho, theta + pi/2) are the same ellipse."""
    rng = np.random.default_rng(12)
    n_tangent = 0
    for i in range(400):
        xc, zc = rng.uniform(-1, 1), rng.uniform(-1, 1)
        a = np.array([xc + rng.uniform(-2, 2), -10.0, zc + rng.uniform(-2, 2)])
        b = np.array([xc + rng.uniform(-2, 2), 30.0, zc + rng.uniform(-2, 2)])
        if i % 4 == 0:  # near tangency: the line passes at distance 1 - tiny
            d = np.array([b[0] - a[0], 0.0, b[2] - a[2]])
            nrm = np.array([-d[2], 0.0, d[0]]) / np.hypot(d[0], d[2])
            off = (1.0 - 10.0 ** rng.uniform(-9, -4)) * nrm
            mid = np.array([xc, 11.0, zc]) + off
            a, b = mid - 0.5 * (b - a), mid + 0.5 * (b - a)
            n_tangent += 1
        c = oracle_mod.cylinder_ray(a, b, xc, 8.0, 14.0, zc, 1.0)
        e = oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, 1.0, 1.0, 0.0)
        assert e == pytest.approx(c, rel=1e-12, abs=1e-12)
        if c > 1e-3:
            th = rng.uniform(-3, 3)
            assert oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, 1.0, 1.0, th) == pytest.approx(c, rel=1e-9)
        xr, zr, th = rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.2), rng.uniform(-3, 3)
        e1 = oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, xr, zr, th)
        e2 = oracle_mod.ellcyl_ray(a, b, xc, 8.0, 14.0, zc, zr, xr, th + math.pi / 2)
        if min(e1, e2) > 1e-3 or max(e1, e2) == 0.0:
            assert e1 == pytest.approx(e2, rel=1e-9)
    assert n_tangent == 100


def test_ellcyl_through_both_caps(oracle_mod):
    """A ray through both end faces of a turned ellipse: (y2 - y1)/|k_y|."""
    a, b = np.array([0.1, -20.0, -0.2]), np.array([0.5, 40.0, 0.3])
    k = (b - a) / np.linalg.norm(b - a)
    got = oracle_mod.ellcyl_ray(a, b, 0.3, 10.0, 15.0, 0.05, 0.8, 0.3, 0.4)
    assert got == pytest.approx(5.0 / k[1], rel=1e-13)
    # the same line misses the ellipse turned by pi/2 (minor axis across it)
    a2, b2 = np.array([0.85, -20.0, 0.0]), np.array([0.85, 40.0, 0.0])
    assert oracle_mod.ellcyl_ray(a2, b2, 0.3, 10.0, 15.0, 0.0, 0.8, 0.3, 0.0) > 0.0
    assert oracle_mod.ellcyl_ray(a2, b2, 0.3, 10.0, 15.0, 0.0, 0.8, 0.3, math.pi / 2) == 0.0


def test_kn_textbook(oracle_mod):
    """Factorised KN of scatter.cu:962-977 equals the textbook form
    P^2 (P + 1/P - sin^2 theta), P = E'/E0, i.e. 2/r_e^2 dsigma/dOmega."""
    for e0 in (70.0, 140.5, 364.5, 511.0):
        for th in np.linspace(0.0, math.pi, 37):
            P = 1.0 / (1.0 + e0 / 511.0 * (1.0 - math.cos(th)))
            text = P * P * (P + 1.0 / P - math.sin(th) ** 2)
            assert oracle_mod.kn(th, e0) == pytest.approx(text, rel=1e-13)


def test_kn_normaliser_golden(oracle_mod):
    """The [0, pi] normaliser with the reference's float step and end
    (0.01f, (float)M_PI; scatter.cu:1169) against an mpmath evaluation."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    step = mp.mpf(float(np.float32(0.01)))
    hi = mp.mpf(float(np.float32(math.pi)))
    n = int(hi / step)
    e0 = mp.mpf(140.5)
    total = mp.mpf(0)
    for i in range(n + 1):
        th = i * step
        P = 1 / (1 + e0 / 511 * (1 - mp.cos(th)))
        total += P * P * (P + 1 / P - mp.sin(th) ** 2) * mp.sin(th) * step
    got = oracle_mod.kn_sum(140.5, 0.0, float(np.float32(math.pi)), float(np.float32(0.01)))
    assert got == pytest.approx(float(total), rel=1e-13)
    # the rectangle sum approximates the closed-form integral of KN sin
    a = 140.5 / 511.0
    closed = ((1 + a) / a ** 2 * (2 * (1 + a) / (1 + 2 * a) - math.log(1 + 2 * a) / a)
              + math.log(1 + 2 * a) / (2 * a) - (1 + 3 * a) / (1 + 2 * a) ** 2)
    assert got == pytest.approx(2 * closed, rel=2e-3)


def test_kn_sum_step_count(oracle_mod):
    """Inclusive rule with n = (int)((hi-lo)/step): n + 1 terms."""
    e0, step = 140.5, 0.01
    lo, hi = 0.3, 0.3 + 5 * step + 0.004
    expect = sum(oracle_mod.kn(lo + i * step, e0) * math.sin(lo + i * step) * step for i in range(6))
    assert oracle_mod.kn_sum(e0, lo, hi, step) == pytest.approx(expect, rel=1e-14)
    # an empty range still counts the first term
    assert oracle_mod.kn_sum(e0, lo, lo, step) == pytest.approx(
        oracle_mod.kn(lo, e0) * math.sin(lo) * step, rel=1e-14)


def test_scatter_energy_limits(oracle_mod):
    e0 = 140.5
    assert oracle_mod.scatter_energy(0.0, e0) == e0
    assert oracle_mod.scatter_energy(math.pi, e0) == pytest.approx(e0 / (1 + 2 * e0 / 511.0), rel=1e-15)
    # Compton formula in wavelength form: 1/E' - 1/E0 = (1 - cos)/511
    th = 1.1
    assert 1 / oracle_mod.scatter_energy(th, e0) - 1 / e0 == pytest.approx((1 - math.cos(th)) / 511.0, rel=1e-12)


def test_cone_angle(oracle_mod):
    assert oracle_mod.cone_angle([0, 0, 0], [0, 1, 0], [0, 2, 0]) == 0.0
    assert oracle_mod.cone_angle([0, 0, 0], [0, 1, 0], [1, 1, 0]) == pytest.approx(math.pi / 2)
    assert oracle_mod.cone_angle([0, 0, 0], [0, 1, 0], [0, 0, 0]) == pytest.approx(math.pi)
    rng = np.random.default_rng(4)
    for _ in range(100):
        s, a, d = rng.normal(size=(3, 3))
        u, w = a - s, d - a
        ref = math.atan2(np.linalg.norm(np.cross(u, w)), u @ w)
        assert oracle_mod.cone_angle(s, a, d) == pytest.approx(ref, abs=1e-12)


def test_gauss_window(oracle_mod):
    """0.5 [erf(z2) - erf(z1)], sigma = res/2.35482 * E (scatter.cu:1025-1033),
    against scipy's normal CDF."""
    fwhm = float(np.float32(2.35482))
    for e in (110.0, 126.0, 140.5, 150.0):
        for res in (0.05, 0.1, 0.2):
            sigma = res / fwhm * e
            ref = special.ndtr((154.55 - e) / sigma) - special.ndtr((126.45 - e) / sigma)
            assert oracle_mod.gauss_window(e, res, 126.45, 154.55) == pytest.approx(ref, rel=1e-12, abs=1e-15)


def test_mu_tables_loglog_interpolation(oracle_mod, params_mod):
    """The oracle's mu(E) tables: power laws are reproduced exactly, between
    two nodes log mu is linear in log E (the value at the geometric mean of
    the node energies is the geometric mean of the node values), and the end
    segments continue beyond the nodes."""
    o = oracle_mod.Oracle(params_mod.config(1))
    E = 20.0 * 10.0 ** (np.arange(64) / 63.0)
    e0 = 140.5
    o.set_mu_tables(E, 0.18 * (e0 / E) ** 3, 0.07 * (e0 / E), 3.3 * (e0 / E) ** 3, 0.35 * (e0 / E))
    for e in (15.0, 90.6, 126.45, 140.5, 154.55, 230.0):
        assert o.mu_ratio(0, e, e0) == pytest.approx((e0 / e) ** 3, rel=1e-13)
        assert o.mu_ratio(1, e, e0) == pytest.approx(e0 / e, rel=1e-13)
        assert o.mu_ratio(2, e, e0) == pytest.approx((e0 / e) ** 3, rel=1e-13)
    # two nodes, arbitrary values
    En = np.array([100.0, 150.0])
    v = np.array([7.0, 2.0])
    o.set_mu_tables(En, v, v[::-1].copy(), v, v)
    gm = np.sqrt(100.0 * 150.0)
    assert o.mu_ratio(0, gm, 100.0) == pytest.approx(np.sqrt(7.0 * 2.0) / 7.0, rel=1e-13)
    assert o.mu_ratio(1, gm, 150.0) == pytest.approx(np.sqrt(7.0 * 2.0) / 7.0, rel=1e-13)
    slope = np.log(2.0 / 7.0) / np.log(1.5)
    assert o.mu_ratio(0, 50.0, 100.0) == pytest.approx(0.5 ** slope, rel=1e-13)
    with pytest.raises(ValueError):
        o.set_mu_tables(En[::-1].copy(), v, v, v, v)
    o.set_mu_tables(None)
    assert o.mu_ratio(0, 100.0, 140.5) == pytest.approx((1.405) ** 3, rel=1e-15)


def _sampled_length(a, b, inside, n=200001):
    t = (np.arange(n) + 0.5) / n
    pts = np.asarray(a)[None, :] + t[:, None] * (np.asarray(b) - np.asarray(a))[None, :]
    return inside(pts).mean() * np.linalg.norm(np.asarray(b) - np.asarray(a))


def test_seg_box_and_seg_cyl_against_sampling(oracle_mod):
    """The grid-scatterer extension's plain intersection lengths against
    dense sampling of random segments (start and end inside or outside)."""
    rng = np.random.default_rng(17)
    lo, hi = np.array([-2.0, 1.0, -1.5]), np.array([3.0, 4.0, 2.5])
    for _ in range(40):
        a = rng.uniform(-5, 6, 3)
        b = rng.uniform(-5, 6, 3) if rng.random() < 0.7 else rng.uniform(lo, hi)
        ref = _sampled_length(a, b, lambda x: np.all((x >= lo) & (x <= hi), axis=1))
        assert oracle_mod.seg_box(a, b, lo, hi) == pytest.approx(ref, abs=2e-4 * np.linalg.norm(b - a))
    xc, zc, r, y1, y2 = 0.4, -0.3, 0.8, -1.0, 2.0
    inside_cyl = lambda x: ((x[:, 0] - xc) ** 2 + (x[:, 2] - zc) ** 2 <= r * r) & (x[:, 1] >= y1) & (x[:, 1] <= y2)
    for _ in range(40):
        a = rng.uniform(-2, 2, 3)
        b = a + rng.normal(0, 1, 3) * np.array([0.6, 3.0, 0.6])
        ref = _sampled_length(a, b, inside_cyl)
        assert oracle_mod.seg_cyl(a, b, xc, y1, y2, zc, r) == pytest.approx(ref, abs=2e-4 * np.linalg.norm(b - a))
    # along the axis: the overlap with [y1, y2]
    assert oracle_mod.seg_cyl([xc, -5.0, zc], [xc, 1.0, zc], xc, y1, y2, zc, r) == pytest.approx(2.0)


def test_seg_ell_against_sampling(oracle_mod):
    """The elliptic-cylinder segment length (grid scatterers behind elliptical
    holes) against dense sampling of random segments through a turned
    ellipse; a circle's semi-axes give the cylinder's length at any turn."""
    rng = np.random.default_rng(18)
    xc, zc, y1, y2 = 0.4, -0.3, -1.0, 2.0
    for xr, zr, th in ((0.9, 0.35, 0.0), (0.5, 1.1, 0.7), (0.3, 0.8, -2.2)):
        def inside(x, xr=xr, zr=zr, th=th):
            dx, dz = x[:, 0] - xc, x[:, 2] - zc
            u = dx * np.cos(th) + dz * np.sin(th)
            w = -dx * np.sin(th) + dz * np.cos(th)
            return (u * u / xr ** 2 + w * w / zr ** 2 <= 1.0) & (x[:, 1] >= y1) & (x[:, 1] <= y2)
        for _ in range(30):
            a = rng.uniform(-2, 2, 3)
            b = a + rng.normal(0, 1, 3) * np.array([0.8, 3.0, 0.8])
            ref = _sampled_length(a, b, inside)
            got = oracle_mod.seg_ell(a, b, xc, y1, y2, zc, xr, zr, th)
            assert got == pytest.approx(ref, abs=2e-4 * np.linalg.norm(b - a)), (xr, zr, th)
    for _ in range(20):
        a = rng.uniform(-2, 2, 3)
        b = a + rng.normal(0, 1, 3) * np.array([0.6, 3.0, 0.6])
        want = oracle_mod.seg_cyl(a, b, xc, y1, y2, zc, 0.7)
        assert oracle_mod.seg_ell(a, b, xc, y1, y2, zc, 0.7, 0.7, rng.uniform(-3, 3)) == pytest.approx(
            want, rel=1e-12, abs=1e-12)


def test_coll_grid_points_leave_out_elliptical_holes(oracle_mod, params_mod):
    """Grid scatterers behind elliptical holes: no grid point lies inside a
    turned ellipse (within the hole's y span), the points between an ellipse
    and its bounding circle are kept, and every other cell centre of the plate
    is a point."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    n_h = p.n_holes_layer0
    shapes = np.zeros((n_h, 3), np.float32)
    rng = np.random.default_rng(23)
    shapes[:, 0] = rng.uniform(0.6, 1.4, n_h)
    shapes[:, 1] = rng.uniform(0.3, 0.7, n_h)
    shapes[:, 2] = rng.uniform(-np.pi, np.pi, n_h)
    o.set_hole_shapes(shapes)
    pitch = 0.25
    pts, _ = o.coll_grid_points(pitch)
    w, t, h = (float(p.coll[k]) for k in (11, 12, 13))
    F = float(p.img[11])
    xs = -w / 2 + (np.arange(round(w / pitch)) + 0.5) * pitch
    ys = -t / 2 + F + float(p.coll[14]) + (np.arange(round(t / pitch)) + 0.5) * pitch
    zs = -h / 2 + (np.arange(round(h / pitch)) + 0.5) * pitch
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    allp = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    inside = np.zeros(len(allp), bool)
    ring = np.zeros(len(allp), bool)
    for i in range(n_h):
        hp = p.coll[100 + 9 * i:100 + 9 * i + 5].astype(np.float64)
        xr, zr, th = (float(v) for v in shapes[i])
        dx, dz = allp[:, 0] - hp[0], allp[:, 2] - hp[3]
        u = dx * np.cos(th) + dz * np.sin(th)
        v = -dx * np.sin(th) + dz * np.cos(th)
        iny = (allp[:, 1] >= hp[1] + F) & (allp[:, 1] <= hp[2] + F)
        ins = (u * u / (xr * xr) + v * v / (zr * zr) < 1.0) & iny
        inside |= ins
        ring |= ~ins & iny & (dx * dx + dz * dz < max(xr, zr) ** 2)
    assert ring.sum() > 100
    keep = allp[~inside]
    assert len(pts) == len(keep)
    assert np.allclose(pts, keep, rtol=0, atol=1e-9)


def test_coll_grid_points_and_weight(oracle_mod, params_mod):
    """Grid over the layer-0 plate: the cells tile the plate less the holes,
    and the first-leg weight is dv mu_c / (4 pi L^2) exp(-mu_t l_W) with l_W
    the tungsten path (here a point behind a hole-free stretch of plate)."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    pts, wts = o.coll_grid_points(1.0)
    dv = wts[0] / float(p.coll[17])
    assert np.all(wts == wts[0])
    w, t, h = p.coll[11], p.coll[12], p.coll[13]
    holes = sum(np.pi * p.coll[100 + 9 * i + 4] ** 2 * (p.coll[100 + 9 * i + 2] - p.coll[100 + 9 * i + 1])
                for i in range(int(p.coll[10])))
    assert len(pts) * dv == pytest.approx(w * t * h - holes, rel=2e-3)
    F = p.img[11]
    y_front = -t / 2 + F + p.coll[14]
    # a source straight in front of a point s, the path avoiding every hole
    hx = np.array([p.coll[100 + 9 * i] for i in range(int(p.coll[10]))])
    hz = np.array([p.coll[100 + 9 * i + 3] for i in range(int(p.coll[10]))])
    cand = [q for q in pts if np.min((hx - q[0]) ** 2 + (hz - q[2]) ** 2) > 4.0]
    s = cand[len(cand) // 2]
    src = np.array([s[0], y_front - 20.0, s[2]])
    depth = s[1] - y_front
    L = np.linalg.norm(s - src)
    ref = dv * p.coll[17] / (4 * np.pi * L * L) * np.exp(-p.coll[15] * depth)
    assert o.coll_grid_weight(src, s, wts[0]) == pytest.approx(ref, rel=1e-12)


def test_coll_grid_points_of_several_layers(oracle_mod, params_mod):
    """Three plates: each layer's points tile its plate less its own holes,
    carry that layer's dv mu_c, and a mask of several layers lists the
    layers' points in increasing layer order."""
    from conftest import small_geometry
    p = small_geometry(params_mod, layers=3)
    for m in range(3):
        p.coll[(m + 1) * 10 + 7] = 0.35 + 0.1 * m  # a Compton mu of each layer's own
    o = oracle_mod.Oracle(p)
    per = [o.coll_grid_points(0.5, layers=1 << m) for m in range(3)]
    first = 0
    for m, (pts, wts) in enumerate(per):
        L = p.coll[(m + 1) * 10:(m + 1) * 10 + 8]
        nh = int(L[0])
        holes = sum(np.pi * p.coll[100 + 9 * i + 4] ** 2 * (p.coll[100 + 9 * i + 2] - p.coll[100 + 9 * i + 1])
                    for i in range(first, first + nh))
        first += nh
        n = [max(1, int(np.floor(e / 0.5 + 0.5))) for e in (L[1], L[2], L[3])]
        dv = (L[1] / n[0]) * (L[2] / n[1]) * (L[3] / n[2])
        assert len(pts) > 0 and np.allclose(wts, dv * L[7], rtol=1e-12, atol=0)
        assert len(pts) * dv == pytest.approx(L[1] * L[2] * L[3] - holes, rel=0.05)
        y0 = -L[2] / 2 + p.img[11] + L[4]
        assert np.all((pts[:, 1] > y0) & (pts[:, 1] < y0 + L[2]))
    allp, allw = o.coll_grid_points(0.5, layers=0b111)
    np.testing.assert_array_equal(allp, np.concatenate([q for q, _ in per]))
    np.testing.assert_array_equal(allw, np.concatenate([w for _, w in per]))


def test_crystal_grid_weight_closed_form(oracle_mod, params_mod):
    """First-leg weight of a crystal grid point with no collimator and no
    other bar on the way: dv mu_c / (4 pi L^2) exp(-mu_t l), l the path in the
    bar from its front face to the point (the voxel straight in front)."""
    p = params_mod.config(1)
    p.coll[0] = 0  # no collimator layer
    b = 9
    xb, zb = float(p.det[12 * b + 1]), float(p.det[12 * b + 3])
    yb = float(p.det[12 * b + 2]) + float(p.img[11])
    w, t, h = (float(x) for x in p.det[12 * b + 4:12 * b + 7])
    mut, muc = float(p.det[12 * b + 7]), float(p.det[12 * b + 9])
    grid = (2, 4, 2)
    # grid column (a, c) = (0, 1) of bar b: x = xb - w/4, z = zb + h/4; put the
    # voxel row z there and pick the voxel column at that x
    p.img[10] = zb + h / 4
    o = oracle_mod.Oracle(p)
    ix = int(round(xb - w / 4 + 15.5))
    assert abs(ix - 15.5 - (xb - w / 4)) < 1e-9
    v = ix + 16 * 32  # iy = 16, y = 0.5
    pv = np.array([ix - 15.5, 0.5, zb + h / 4])
    dv = w * t * h / 16
    for by in range(4):
        cell = (1 * 2 + 0) * 4 + by
        s = np.array([xb - w / 4, yb - t / 2 + (by + 0.5) * t / 4, zb + h / 4])
        L = np.linalg.norm(s - pv)
        ref = dv * muc / (4 * np.pi * L * L) * np.exp(-mut * (s[1] - (yb - t / 2)))
        assert o.crystal_grid_weight(b, v, grid, cell) == pytest.approx(ref, rel=1e-12)
