"""Independent pins of the oracle's scatter element and of the attenuation
through turned bars.

The oracle restates the reference's scatter element
(ScatterGen_RayTracing_CircularHole/scatter.cu:1226-1427) and PE element
(PEGen_RayTracing_CircularHole/PESysMatGen.cu:1193-1363) in FP64 with
bar-frame slab chords, acos cone angles and the factorised Klein-Nishina
form.  The checks here reach the same quantities by other routes, so that a
dropped term, a swapped energy scaling or a wrong frame fails one of them:

* single (voxel, scatter point, detection cell) terms evaluated in mpmath at
  30 digits: cone angles from atan2(|u x w|, u . w), the eight corner angles
  and every chord from plane clipping of the world-frame segment against the
  turned box (no bar-frame transform), Klein-Nishina in its textbook form
  P^2 (P + 1/P - sin^2 theta), the window probability from the normal CDF,
  E' from the wavelength shift;
* derivatives that the model fixes: with one attenuating bar k on the second
  leg, d ln(element) / d mu_pe(k) = -(E0/E')^3 l_k and
  d ln(element) / d mu_c(k) = -(E0/E') l_k, and the bar's mu_total parameter
  does not enter (scatter.cu:1371-1375); in the PE element
  d ln A / d mu_total(k) = -l_k and mu_pe(k) does not enter
  (PESysMatGen.cu:1290-1325);
* invariance of both elements when voxel and bars are turned together about
  the y axis (bar angle beta + gamma, positions turned by gamma), and the
  quarter-turned bar equal to the unturned bar with width and height swapped.
"""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 30

F = 30.0
E0 = 140.5
FWHM = float(np.float32(2.35482))      # scatter.cu:1026 (float literal)
KN_STEP = float(np.float32(0.01))      # scatter.cu:1338 (0.01 passed as float)
KN_PI = float(np.float32(math.pi))     # scatter.cu:1169 (M_PI passed as float)
BAR = dict(w=2.5, t=8.0, h=2.5)


# ---- independent model --------------------------------------------------------------

def mpv(x):
    return [mp.mpf(float(c)) for c in x]


def sub(a, b):
    return [a[i] - b[i] for i in range(3)]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def norm(a):
    return mp.sqrt(dot(a, a))


def angle(u, w):
    """Angle between u and w from atan2 of the cross and dot products."""
    return mp.atan2(norm(cross(u, w)), dot(u, w))


class Box:
    """A bar or sub-cell in world coordinates: centre, half extents along the
    turned axes ex = (cos b, 0, sin b), ey = y, ez = (-sin b, 0, cos b)."""

    def __init__(self, centre, half, beta):
        self.c = mpv(centre)
        self.half = mpv(half)
        b = mp.mpf(float(beta))
        self.axes = [[mp.cos(b), mp.mpf(0), mp.sin(b)], [mp.mpf(0), mp.mpf(1), mp.mpf(0)],
                     [-mp.sin(b), mp.mpf(0), mp.cos(b)]]

    def corners(self):
        out = []
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    s = (sx, sy, sz)
                    out.append([self.c[i] + sum(s[k] * self.half[k] * self.axes[k][i] for k in range(3))
                                for i in range(3)])
        return out

    def clip(self, a, b):
        """Parameter interval [t0, t1] of the line a + t (b - a) inside the box."""
        t0, t1 = mp.mpf("-inf"), mp.mpf("inf")
        d = sub(b, a)
        r = sub(a, self.c)
        for k in range(3):
            s0 = dot(r, self.axes[k])
            ds = dot(d, self.axes[k])
            lo, hi = -self.half[k] - s0, self.half[k] - s0
            if ds == 0:
                if lo > 0 or hi < 0:
                    return None
                continue
            ta, tb = lo / ds, hi / ds
            t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
        return (t0, t1) if t1 > t0 else None

    def chord(self, a, b):
        """Length of the segment a -> b inside the box."""
        iv = self.clip(a, b)
        if iv is None:
            return mp.mpf(0)
        t0, t1 = max(iv[0], 0), min(iv[1], 1)
        return max(t1 - t0, 0) * norm(sub(b, a))

    def contains(self, p):
        r = sub(p, self.c)
        return all(abs(dot(r, self.axes[k])) < self.half[k] for k in range(3))


def bar_box(x, y, z, beta, w=BAR["w"], t=BAR["t"], h=BAR["h"]):
    return Box((x, y + F, z), (w / 2, t / 2, h / 2), beta)


def cells(bar, x, y, z, beta, split):
    """Sub-cells of a bar: the (a, b, c)-th of sx x sy x sz equal slices."""
    w, t, h = bar["w"], bar["t"], bar["h"]
    sx, sy, sz = split
    b = mp.mpf(float(beta))
    out = []
    for c in range(sz):
        for a in range(sx):
            for bb in range(sy):
                lx = -mp.mpf(w) / 2 + (a + mp.mpf(0.5)) * mp.mpf(w) / sx
                ly = -mp.mpf(t) / 2 + (bb + mp.mpf(0.5)) * mp.mpf(t) / sy
                lz = -mp.mpf(h) / 2 + (c + mp.mpf(0.5)) * mp.mpf(h) / sz
                cx = mp.mpf(x) + lx * mp.cos(b) - lz * mp.sin(b)
                cz = mp.mpf(z) + lx * mp.sin(b) + lz * mp.cos(b)
                out.append(Box((cx, mp.mpf(y) + F + ly, cz), (mp.mpf(w) / sx / 2, mp.mpf(t) / sy / 2,
                                                        mp.mpf(h) / sz / 2), beta))
    return out


def kn_textbook(theta):
    """dsigma/dOmega without r_e^2 / 2: P^2 (P + 1/P - sin^2 theta)."""
    P = 1 / (1 + mp.mpf(E0) / 511 * (1 - mp.cos(theta)))
    return P * P * (P + 1 / P - mp.sin(theta) ** 2)


def kn_rect(lo, hi):
    """The reference's inclusive rectangle rule on lo, lo + step, ...
    (scatter.cu:980-993)."""
    q = (hi - lo) / KN_STEP
    n = int(mp.floor(q))
    assert abs(q - mp.nint(q)) > 1e-9, "node count too close to an integer for a pin"
    return mp.fsum(kn_textbook(lo + i * mp.mpf(KN_STEP)) * mp.sin(lo + i * mp.mpf(KN_STEP)) *
                   KN_STEP for i in range(n + 1))


KN_NORM = None


def kn_norm():
    global KN_NORM
    if KN_NORM is None:
        KN_NORM = kn_rect(mp.mpf(0), mp.mpf(KN_PI))
    return KN_NORM


def term(p, qs, det_bar, cell, res, window, others, mu_d, diag=None):
    """One (voxel p, scatter point qs, detection cell) term without its
    first-leg weight.  others: [(Box, mu_pe, mu_c)] flagged on the second leg;
    mu_d = (mu_pe, mu_c) of the detection bar.  Returns (value, info)."""
    q = cell.c
    u, w = sub(qs, p), sub(q, qs)
    theta = angle(u, w)
    # E' from the wavelength shift: 1/E' - 1/E0 = (1 - cos theta) / 511 keV
    ep = 1 / (1 / mp.mpf(E0) + (1 - mp.cos(theta)) / 511)
    resp = mp.mpf(res) * mp.sqrt(ep / E0)
    lo, hi = mp.mpf(window[0]), mp.mpf(window[1])
    info = dict(theta=theta, ep=ep)
    if ep * (1 + 2 * resp / FWHM) <= lo or ep * (1 - 2 * resp / FWHM) >= hi:
        info["cut"] = True
        return mp.mpf(0), info
    sigma = resp / FWHM * ep
    p_win = mp.ncdf((hi - ep) / sigma) - mp.ncdf((lo - ep) / sigma)
    L = norm(w)
    half = cell.half
    rdet = mp.sqrt((2 * half[0]) ** 2 + (2 * half[1]) ** 2 + (2 * half[2]) ** 2) / 2
    sin_t = norm(cross(u, w)) / (norm(u) * L)
    if 2 * L * sin_t <= rdet:
        phi = 2 * mp.pi
        info["phi_branch"] = "2pi"
    else:
        phi = 4 * mp.asin(rdet / (2 * L * sin_t))
        info["phi_branch"] = "asin"
    ths = [angle(u, sub(c, qs)) for c in cell.corners()]
    tmin, tmax = min(ths), max(ths)
    # the voxel -> scatter point line, 1000 mm past the scatter point, through the cell
    ext = [qs[i] + 1000 * u[i] / norm(u) for i in range(3)]
    if cell.chord(p, ext) > mp.mpf(float(np.float32(0.001))):
        tmin = mp.mpf(0)
        info["through_cell"] = True
    ratio = kn_rect(tmin, tmax) / kn_norm()
    e3, e1 = (E0 / ep) ** 3, E0 / ep
    att = mp.mpf(0)
    for box, mpe, mc in others:
        att += box.chord(qs, q) * (mp.mpf(mpe) * e3 + mp.mpf(mc) * e1)
    mupe_d, mut_d = mp.mpf(mu_d[0]) * e3, mp.mpf(mu_d[0]) * e3 + mp.mpf(mu_d[1]) * e1
    # path in the detection bar before the cell, then the cell's chord along the line
    iv_bar = det_bar.clip(qs, q)
    iv_cell = cell.clip(qs, q)
    l_bar = (1 - iv_bar[0]) * L
    l_cell = (1 - iv_cell[0]) * L
    att += (l_bar - l_cell) * mut_d
    ext2 = [q[i] + 1000 * w[i] / L for i in range(3)]
    la = cell.chord(qs, ext2)
    val = phi / (2 * mp.pi) * ratio * p_win * mp.exp(-att) * (1 - mp.exp(-la * mut_d)) * mupe_d / mut_d
    info.update(phi=phi, ratio=ratio, p_win=p_win, e3=e3, e1=e1, att=att)
    return val, info


# ---- geometry builders ----------------------------------------------------------------

def scatter_geometry(params_mod, bars, voxel, window=(126.45, 154.55)):
    """bars: [(x, y, z, beta, mu_t, mu_pe, mu_c)], 1x1x1 image at voxel."""
    p = params_mod.Params.empty()
    p.coll[0] = 1
    params_mod.set_layer(p, 0, 0, 300.0, 5.0, 300.0, 0.0, 0.0, 0.0, 0.0)
    p.det[0] = len(bars)
    for k, (x, y, z, beta, mt, mpe, mc) in enumerate(bars):
        params_mod.set_bar(p, k, x, y, z, BAR["w"], BAR["t"], BAR["h"], mt, mpe, mc, 0.1, beta)
    params_mod.set_image(p, 1, 1, 1, 1.0, 1.0, 1.0, shift=voxel, fov2coll=F)
    params_mod.set_physics(p, lo=window[0], hi=window[1])
    return p


def f32(x):
    return float(np.float32(x))


def model_crystal_element(p, d, split, pe_w):
    """Crystal scatter element (scatter.cu:1077-1437) from the model: every
    bar s != d at its centre, first-leg weight pe_w[s] mu_c(d) / mu_pe(d),
    second leg through every other bar the bitmap flags (flags: dict
    (s, k) -> bool)."""
    det = p.det
    n = int(det[0])
    vox = mpv(p.img[8:11])
    bars = [(float(det[12 * k + 1]), float(det[12 * k + 2]), float(det[12 * k + 3]),
             float(det[12 * k + 11]), float(det[12 * k + 7]), float(det[12 * k + 8]),
             float(det[12 * k + 9])) for k in range(n)]
    xd, yd, zd, bd, _, mpe_d, mc_d = bars[d]
    dbox = bar_box(xd, yd, zd, bd)
    total = mp.mpf(0)
    infos = []
    for s in range(n):
        if s == d or pe_w[s] == 0:
            continue
        qs = mpv((bars[s][0], bars[s][1] + F, bars[s][2]))
        others = [(bar_box(*bars[k][:4]), bars[k][5], bars[k][6]) for k in range(n)
                  if k not in (s, d)]
        wgt = mp.mpf(float(np.float32(pe_w[s]))) * mp.mpf(mc_d) / mp.mpf(mpe_d)
        for cell in cells(BAR, xd, yd, zd, bd, split):
            val, info = term(vox, qs, dbox, cell, float(det[12 * d + 10]),
                             (float(p.phys[5]), float(p.phys[6])), others, (mpe_d, mc_d))
            total += wgt * val
            infos.append(info)
    return total, infos


def oracle_crystal_element(oracle_mod, p, d, split, pe_w):
    o = oracle_mod.Oracle(p)
    bm = o.bitmap_crystal()
    pe = np.asarray(pe_w, np.float32).reshape(-1, 1)
    return float(o.crystal_scatter_pairs(pe, bm, [d], [0], sub=split)[0]), o, bm


# ---- single-term evaluations ----------------------------------------------------------

def test_scatter_term_asin_branch_with_corner_range(oracle_mod, params_mod):
    """Scatter bar in front, turned detection bar behind and to the side: the
    phi range takes the asin branch and theta_min comes from the corners.
    Two detection-bar splits (the second exercises the path in the bar before
    the sub-cell)."""
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (5.0, 26.5, 3.0, 0.35, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, bars, voxel=(-1.5, 0.5, 0.75))
    for split in ((1, 1, 1), (2, 3, 2)):
        want, infos = model_crystal_element(p, 1, split, [1e-3, 0.0])
        got, _, _ = oracle_crystal_element(oracle_mod, p, 1, split, [1e-3, 0.0])
        assert all(i.get("phi_branch") == "asin" for i in infos if not i.get("cut"))
        assert not any(i.get("through_cell") for i in infos)
        assert want > 0
        assert got == pytest.approx(float(want), rel=1e-11), split


def test_scatter_term_2pi_branch_through_cell(oracle_mod, params_mod):
    """Voxel, scatter bar and detection bar nearly in line: 2 L sin(theta) <=
    R_cell gives phi = 2 pi, and the voxel -> scatter point line crosses the
    detection cell, so theta_min = 0 (scatter.cu:1281-1283, 1331-1334)."""
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (1.2, 26.0, 0.0, 0.0, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, bars, voxel=(-3.0, 0.0, 0.0))
    want, infos = model_crystal_element(p, 1, (1, 1, 1), [1e-3, 0.0])
    got, _, _ = oracle_crystal_element(oracle_mod, p, 1, (1, 1, 1), [1e-3, 0.0])
    (info,) = infos
    assert info["phi_branch"] == "2pi" and info.get("through_cell")
    assert got == pytest.approx(float(want), rel=1e-11)


def test_scatter_term_window_cut(oracle_mod, params_mod):
    """A detection bar at a right angle to the incoming direction: E' is far
    below the window and the 2-sigma early-out drops the term, in the model
    and in the oracle."""
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (12.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, bars, voxel=(0.0, 0.0, 0.0))
    want, infos = model_crystal_element(p, 1, (1, 1, 1), [1e-3, 0.0])
    got, _, _ = oracle_crystal_element(oracle_mod, p, 1, (1, 1, 1), [1e-3, 0.0])
    assert infos[0].get("cut") and want == 0 and got == 0.0


def test_scatter_term_between_one_and_two_sigma(oracle_mod, params_mod):
    """A scattering angle of 69.3 degrees puts E' ~ 119.3 keV: E' + 2 sigma
    reaches the window's lower edge (126.45 keV) but E' + 1 sigma does not.
    The reference's early-out uses 2 sigma (scatter.cu:1263), so the term is
    kept with its small window probability."""
    th = math.acos(1 - 511.0 * (1 / 119.3 - 1 / E0))
    xd, yd = 12.0 * math.sin(th), 18.0 + 12.0 * math.cos(th)
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07), (xd, yd, 0.0, 0.0, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, bars, voxel=(0.0, 0.0, 0.0))
    want, infos = model_crystal_element(p, 1, (1, 1, 1), [1e-3, 0.0])
    got, _, _ = oracle_crystal_element(oracle_mod, p, 1, (1, 1, 1), [1e-3, 0.0])
    ep = infos[0]["ep"]
    resp = f32(0.1) * mp.sqrt(ep / E0)
    lo = f32(126.45)
    assert ep * (1 + resp / FWHM) <= lo < ep * (1 + 2 * resp / FWHM)
    assert want > 0
    assert got == pytest.approx(float(want), rel=1e-11)


def test_scatter_weight_uses_detection_bar_coefficients(oracle_mod, params_mod):
    """The first-leg weight is PE[s][v] mu_c(d) / mu_pe(d), with the
    detection bar's coefficients (scatter.cu:1219): doubling mu_c of the
    scatter bar changes nothing; changing mu_c(d) changes the weight and the
    absorption as the model says."""
    base = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (5.0, 26.5, 3.0, 0.35, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, base, voxel=(-1.5, 0.5, 0.75))
    a, _, _ = oracle_crystal_element(oracle_mod, p, 1, (1, 1, 1), [1e-3, 0.0])
    q = p.copy()
    q.det[12 * 0 + 9] *= 2.0
    b, _, _ = oracle_crystal_element(oracle_mod, q, 1, (1, 1, 1), [1e-3, 0.0])
    assert a > 0 and b == a
    r = p.copy()
    r.det[12 * 1 + 9] = 0.11
    want, _ = model_crystal_element(r, 1, (1, 1, 1), [1e-3, 0.0])
    c, _, _ = oracle_crystal_element(oracle_mod, r, 1, (1, 1, 1), [1e-3, 0.0])
    assert c == pytest.approx(float(want), rel=1e-11) and abs(c / a - 1) > 0.1


# ---- second-leg attenuation: energy scaling per process ------------------------------

def three_bar_scatter(params_mod, mu_k, beta_k=0.6):
    """Scatter bar 0, turned attenuating bar 2 between it and detection bar 1."""
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (4.0, 34.0, 2.0, 0.2, 0.25, 0.18, 0.07),
            (1.9, 26.0, 1.0, beta_k, *mu_k)]
    return scatter_geometry(params_mod, bars, voxel=(-1.0, 0.0, 0.5))


def test_second_leg_slopes_per_process(oracle_mod, params_mod):
    """One attenuating bar k: ln(element) is linear in its coefficients with
    slopes -(E0/E')^3 l_k (photoelectric) and -(E0/E') l_k (Compton), l_k its
    chord on the scatter point -> cell centre segment; its mu_total
    parameter does not enter (scatter.cu:1371-1375)."""
    pe_w = [1e-3, 0.0, 0.0]
    p0 = three_bar_scatter(params_mod, (0.25, 0.18, 0.07))
    o = oracle_mod.Oracle(p0)
    assert o.crystal_pred(0, 1, 2) == 1
    g0, _, _ = oracle_crystal_element(oracle_mod, p0, 1, (1, 1, 1), pe_w)
    _, infos = model_crystal_element(p0, 1, (1, 1, 1), pe_w)
    ep = infos[0]["ep"]
    kbox = bar_box(f32(1.9), 26.0, 1.0, f32(0.6))
    dq = cells(BAR, 4.0, 34.0, 2.0, f32(0.2), (1, 1, 1))[0].c
    lk = float(kbox.chord(mpv((0.0, 18.0 + F, 0.0)), dq))
    assert lk > 1.0  # the segment crosses the bar, not a corner
    e0ep = float(E0 / ep)

    def slope(idx, delta):
        q = p0.copy()
        q.det[12 * 2 + idx] = np.float32(q.det[12 * 2 + idx] + delta)
        g1, _, _ = oracle_crystal_element(oracle_mod, q, 1, (1, 1, 1), pe_w)
        dmu = float(q.det[12 * 2 + idx]) - float(p0.det[12 * 2 + idx])
        return (math.log(g1) - math.log(g0)) / dmu

    assert slope(8, 0.05) == pytest.approx(-e0ep ** 3 * lk, rel=1e-9)   # mu_pe
    assert slope(9, 0.05) == pytest.approx(-e0ep * lk, rel=1e-9)        # mu_c
    q = p0.copy()
    q.det[12 * 2 + 7] = 2.0                                              # mu_total
    g2, _, _ = oracle_crystal_element(oracle_mod, q, 1, (1, 1, 1), pe_w)
    assert g2 == g0
    # and the whole element, with the turned bar on the second leg, from the model
    want, _ = model_crystal_element(p0, 1, (1, 1, 1), pe_w)
    assert g0 == pytest.approx(float(want), rel=1e-11)


def test_detection_bar_absorption_uses_scaled_coefficients(oracle_mod, params_mod):
    """The detection bar's absorption (1 - exp(-l mu_t')) mu_pe'/mu_t' uses
    mu_pe (E0/E')^3 + mu_c (E0/E'), not its mu_total parameter
    (scatter.cu:1412-1427): changing mu_total of bar d leaves the element
    unchanged while it stays above the 0.01 gate."""
    bars = [(0.0, 18.0, 0.0, 0.0, 0.25, 0.18, 0.07),
            (5.0, 26.5, 3.0, 0.35, 0.25, 0.18, 0.07)]
    p = scatter_geometry(params_mod, bars, voxel=(-1.5, 0.5, 0.75))
    a, _, _ = oracle_crystal_element(oracle_mod, p, 1, (2, 2, 2), [1e-3, 0.0])
    q = p.copy()
    q.det[12 + 7] = 0.9
    b, _, _ = oracle_crystal_element(oracle_mod, q, 1, (2, 2, 2), [1e-3, 0.0])
    assert a > 0 and a == b
    q.det[12 + 7] = 0.009  # below the reference's 0.01 gate (scatter.cu:1225)
    c, _, _ = oracle_crystal_element(oracle_mod, q, 1, (2, 2, 2), [1e-3, 0.0])
    assert c == 0.0


# ---- turned frames ----------------------------------------------------------------------

def turn(x, z, gamma):
    return x * math.cos(gamma) - z * math.sin(gamma), x * math.sin(gamma) + z * math.cos(gamma)


def test_scatter_element_invariant_under_common_turn(oracle_mod, params_mod):
    """Turning voxel and all bars by gamma about the y axis (positions turned,
    beta -> beta + gamma) leaves the scatter element unchanged: a bar frame
    taken the wrong way round, or corners taken in the wrong frame, breaks
    this."""
    base = [(0.0, 18.0, 0.0, 0.1, 0.25, 0.18, 0.07),
            (4.0, 34.0, 2.0, 0.2, 0.25, 0.18, 0.07),
            (1.9, 26.0, 1.0, 0.6, 0.25, 0.18, 0.07)]
    vox = (-1.0, 0.0, 0.5)
    pe_w = [1e-3, 0.0, 0.0]
    vals = []
    for gamma in (0.0, 0.45, -1.1):
        bars = []
        for (x, y, z, b, *mu) in base:
            xr, zr = turn(x, z, gamma)
            bars.append((xr, y, zr, b + gamma, *mu))
        vr = turn(vox[0], vox[2], gamma)
        p = scatter_geometry(params_mod, bars, voxel=(vr[0], vox[1], vr[1]))
        g, _, _ = oracle_crystal_element(oracle_mod, p, 1, (2, 2, 2), pe_w)
        vals.append(g)
    assert vals[0] > 0
    # positions and angles are stored as float32: agreement to their rounding
    np.testing.assert_allclose(vals[1:], vals[0], rtol=2e-5)


def pe_geometry(params_mod, bars, voxel, mu_coll=0.0):
    """Wide plate of mu_coll without holes, bars [(x, y, z, beta, mu_t,
    mu_pe, mu_c)], one voxel."""
    p = params_mod.Params.empty()
    p.coll[0] = 1
    params_mod.set_layer(p, 0, 0, 400.0, 5.0, 400.0, 0.0, mu_coll, 0.9 * mu_coll, 0.1 * mu_coll)
    p.det[0] = len(bars)
    for k, (x, y, z, beta, mt, mpe, mc) in enumerate(bars):
        params_mod.set_bar(p, k, x, y, z, BAR["w"], BAR["t"], BAR["h"], mt, mpe, mc, 0.1, beta)
    params_mod.set_image(p, 1, 1, 1, 1.0, 1.0, 1.0, shift=voxel, fov2coll=F)
    params_mod.set_physics(p)
    return p


def test_pe_turned_bar_attenuation_slope(oracle_mod, params_mod):
    """PE element with one sub-cell (a single ray from the voxel to the
    centre of turned bar d) and a turned bar k across it:
    d ln A / d mu_total(k) = -l_k, l_k from plane clipping of the world-frame
    ray against the turned box; mu_pe(k) does not enter (PESysMatGen.cu:
    1290-1325)."""
    bars = [(3.0, 30.0, 1.5, 0.3, 0.25, 0.18, 0.07),
            (1.6, 15.0, 0.8, 0.7, 0.25, 0.18, 0.07)]
    vox = (0.0, 0.0, 0.0)
    p0 = pe_geometry(params_mod, bars, vox)
    a0 = oracle_mod.Oracle(p0).pe_element(0, 0, sub=(1, 1, 1))
    kbox = bar_box(f32(1.6), 15.0, f32(0.8), f32(0.7))
    lk = float(kbox.chord(mpv(vox), mpv((3.0, 30.0 + F, 1.5))))
    assert lk > 1.0 and a0 > 0
    q = p0.copy()
    q.det[12 + 7] = np.float32(0.35)
    a1 = oracle_mod.Oracle(q).pe_element(0, 0, sub=(1, 1, 1))
    s = (math.log(a1) - math.log(a0)) / (float(q.det[12 + 7]) - float(p0.det[12 + 7]))
    assert s == pytest.approx(-lk, rel=1e-9)
    r = p0.copy()
    r.det[12 + 8] = np.float32(0.02)
    assert oracle_mod.Oracle(r).pe_element(0, 0, sub=(1, 1, 1)) == a0


def test_pe_solid_angle_face_projection(oracle_mod, params_mod):
    """Single ray (one sub-cell) from a voxel far to the side of a turned
    bar, no attenuation outside the bar: A = Omega (1 - exp(-mu_t l)) mu_pe /
    mu_t with l the chord of the voxel -> centre line through the bar (plane
    clipping) and Omega the largest of the three face terms.  The reference
    projects D on the x and z faces with D turned by +beta,
    |D_x cos b - D_z sin b| and |D_x sin b + D_z cos b| (PESysMatGen.cu:
    1223-1225), which is not the bar-frame projection for beta != 0; here
    the x face dominates, so a bar-frame projection would differ by > 1%."""
    beta = f32(0.5)
    bars = [(0.0, 30.0, 0.0, beta, 0.25, 0.18, 0.07)]
    vox = (-60.0, 0.0, 9.0)
    p = pe_geometry(params_mod, bars, vox)
    got = oracle_mod.Oracle(p).pe_element(0, 0, sub=(1, 1, 1))
    pv = mpv(vox)
    q = mpv((0.0, 30.0 + F, 0.0))
    D = sub(q, pv)
    d = norm(D)
    b = mp.mpf(beta)
    w, t, h = (mp.mpf(BAR[k]) for k in ("w", "t", "h"))

    def omega(px, pz):
        faces = [t * h * abs(px), w * t * abs(pz), w * h * abs(D[1])]
        return max(faces) / (4 * mp.pi * d ** 3), faces

    om_ref, faces = omega(D[0] * mp.cos(b) - D[2] * mp.sin(b), D[0] * mp.sin(b) + D[2] * mp.cos(b))
    om_frame, _ = omega(D[0] * mp.cos(b) + D[2] * mp.sin(b), -D[0] * mp.sin(b) + D[2] * mp.cos(b))
    assert faces[0] == max(faces)
    box = bar_box(0.0, 30.0, 0.0, beta)
    ext = [q[i] + 1000 * D[i] / d for i in range(3)]
    la = box.chord(pv, ext)
    mt, mpe = mp.mpf(f32(0.25)), mp.mpf(f32(0.18))
    want = om_ref * (1 - mp.exp(-la * mt)) * mpe / mt
    assert got == pytest.approx(float(want), rel=1e-11)
    assert abs(om_frame / om_ref - 1) > 0.01


def test_pe_element_invariant_under_common_turn(oracle_mod, params_mod):
    """Turning voxel, detection bar and blocking bar together about the y
    axis leaves the PE element unchanged (plate of zero mu, wider than the
    rays): pins the bar frames of the sub-cell points, the own-bar chords
    and the other-bar chords.  The rays run close to y, so the y face term
    is the solid angle's maximum for every sub-cell (the x and z face terms
    use the reference's +beta projection, pinned separately)."""
    base = [(3.0, 30.0, 1.5, 0.3, 0.25, 0.18, 0.07),
            (1.6, 15.0, 0.8, 0.7, 0.25, 0.18, 0.07)]
    vox = (0.4, 0.0, -0.3)
    vals = []
    for gamma in (0.0, 0.8, -0.5):
        bars = []
        for (x, y, z, b, *mu) in base:
            xr, zr = turn(x, z, gamma)
            bars.append((xr, y, zr, b + gamma, *mu))
        vr = turn(vox[0], vox[2], gamma)
        p = pe_geometry(params_mod, bars, (vr[0], vox[1], vr[1]))
        vals.append(oracle_mod.Oracle(p).pe_element(0, 0, sub=(2, 3, 2)))
    assert vals[0] > 0
    np.testing.assert_allclose(vals[1:], vals[0], rtol=2e-5)


def test_pe_quarter_turned_blocking_bar_is_swapped_bar(oracle_mod, params_mod):
    """A blocking bar of width a and height b turned by pi/2 occupies the
    same box as an unturned bar of width b and height a."""
    det = (3.0, 30.0, 1.5, 0.0, 0.25, 0.18, 0.07)
    p1 = pe_geometry(params_mod, [det, (1.6, 15.0, 0.8, math.pi / 2, 0.25, 0.18, 0.07)], (0.0, 0.0, 0.0))
    params_mod.set_bar(p1, 1, 1.6, 15.0, 0.8, 1.5, 8.0, 3.5, 0.25, 0.18, 0.07, 0.1, math.pi / 2)
    p2 = p1.copy()
    params_mod.set_bar(p2, 1, 1.6, 15.0, 0.8, 3.5, 8.0, 1.5, 0.25, 0.18, 0.07, 0.1, 0.0)
    a1 = oracle_mod.Oracle(p1).pe_element(0, 0, sub=(2, 3, 2))
    a2 = oracle_mod.Oracle(p2).pe_element(0, 0, sub=(2, 3, 2))
    p3 = p2.copy()
    p3.det[12 + 7] = 0.0  # no blocking: the bar must really attenuate
    a3 = oracle_mod.Oracle(p3).pe_element(0, 0, sub=(2, 3, 2))
    assert a1 > 0 and a3 > 1.05 * a2
    assert a1 == pytest.approx(a2, rel=1e-6)


# ---- collimator scatter at a hole centre ---------------------------------------------

def test_collimator_scatter_single_hole_term(oracle_mod, params_mod):
    """One layer-0 hole, one detection bar and a turned attenuating bar: the
    element is the hole weight (pi r^2 + T r) / (4 pi L^2) |dy| / L x
    mu_c / mu_total of layer 0 (r = 0.1, scatter.cu:1702-1716) times the
    model's term from the hole centre (x, (y1 + y2)/2 + F, z)."""
    p = params_mod.Params.empty()
    p.coll[0] = 1
    params_mod.set_layer(p, 0, 1, 60.0, 5.0, 60.0, 0.0, 3.65, 3.30, 0.35)
    params_mod.set_hole(p, 0, 0.7, -2.5, 2.5, -0.4, 0.5)
    bars = [(4.0, 20.0, 2.0, 0.2, 0.25, 0.18, 0.07),
            (2.3, 9.5, 0.8, 0.6, 0.25, 0.18, 0.07)]
    p.det[0] = 2
    for k, (x, y, z, b, mt, mpe, mc) in enumerate(bars):
        params_mod.set_bar(p, k, x, y, z, BAR["w"], BAR["t"], BAR["h"], mt, mpe, mc, 0.1, b)
    vox = (-1.5, 0.0, -1.25)
    params_mod.set_image(p, 1, 1, 1, 1.0, 1.0, 1.0, shift=vox, fov2coll=F)
    params_mod.set_physics(p)
    o = oracle_mod.Oracle(p)
    bm = o.bitmap_collimator()
    assert o.coll_pred(0, 0, 1) == 1
    got = float(o.collimator_scatter_pairs(bm, [0], [0], variant=0, sub=(2, 2, 2))[0])

    pv = mpv(vox)
    qs = mpv((f32(0.7), 0.0 + F, f32(-0.4)))
    L = norm(sub(qs, pv))
    r, T = mp.mpf(f32(0.1)), mp.mpf(5.0)
    wgt = (mp.pi * r * r + T * r) / (4 * mp.pi * L * L) * abs(qs[1] - pv[1]) / L
    wgt *= mp.mpf(f32(0.35)) / mp.mpf(f32(3.65))
    x, y, z, b = (f32(c) for c in bars[0][:4])
    dbox = bar_box(x, y, z, b)
    kbox = bar_box(*(f32(c) for c in bars[1][:4]))
    want = mp.mpf(0)
    for cell in cells(BAR, x, y, z, b, (2, 2, 2)):
        val, _ = term(pv, qs, dbox, cell, f32(0.1), (f32(126.45), f32(154.55)),
                      [(kbox, f32(0.18), f32(0.07))], (f32(0.18), f32(0.07)))
        want += wgt * val
    assert want > 0
    assert got == pytest.approx(float(want), rel=1e-11)
