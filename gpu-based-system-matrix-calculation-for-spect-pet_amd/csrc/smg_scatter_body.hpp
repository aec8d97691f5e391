// Per-lane body of the primary-Compton scatter kernels.
//
// Reference: crystalScatterSysMatCuda and collimatorScatterSysMatCuda
// (ScatterGen_RayTracing_CircularHole/scatter.cu:1077-1437, :1544-1936;
// standalone variant ScatterGen_Collimator/scatter.cu:1056-1435).  One term is
// (voxel v, scatter point s, detection sub-cell of bar d):
//
//   w_s(v) * Phi/(2 pi) * K(theta_min, theta_max) * P_E * exp(-att2) * abs'
//
// The second-leg attenuation runs from s to the sub-cell, so its chords do
// not depend on v; only the energy scaling does, through E'(theta(v)):
//   att2 = (E0/E')^3 * S_pe + (E0/E') * S_c,
//   S_pe = sum_k l_k mu_pe_k + (l1 - l2) mu_pe_d,  S_c likewise with mu_c,
// the same sum the reference forms term by term (scatter.cu:1343-1417).
// S_pe, S_c, the absorption chord and |s - q| are computed once per
// (d, s, sub-cell) by the FP64 pair-terms kernel and read here.
#pragma once

#include <cmath>

#include "smg_pe_body.hpp"

namespace smg {

struct PairTerm {
    double s_pe, s_c, la, len;  // len = |scatter point - sub-cell centre|
    double s_wpe, s_wc;         // collimator material on the second leg (grid scatterers)
};

// ---- collimator material for grid scatterers ----------------------------------------
// (An extension with no reference counterpart.)  Plain intersection lengths of
// a segment with a box and with a finite cylinder along y, without the
// reference chord functions' thresholds, in FP64.

SMG_FN double seg_box_geo(const double a[3], const double b[3], const double lo[3],
                          const double hi[3]) {
    double t0 = 0.0, t1 = 1.0;
    for (int k = 0; k < 3; ++k) {
        const double D = b[k] - a[k];
        if (D == 0.0) {
            if (a[k] < lo[k] || a[k] > hi[k]) return 0.0;
            continue;
        }
        double u = (lo[k] - a[k]) / D, w = (hi[k] - a[k]) / D;
        if (u > w) {
            const double x = u;
            u = w;
            w = x;
        }
        t0 = tmax2(t0, u);
        t1 = tmin2(t1, w);
    }
    if (!(t1 > t0)) return 0.0;
    const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return (t1 - t0) * sqrt(dx * dx + dy * dy + dz * dz);
}

// Fraction of the segment a -> b (its parameter range in [0, 1]) inside the
// cylinder of radius r about the line x = xc, z = zc, for y1 <= y <= y2.
SMG_FN double seg_cyl_frac(const double a[3], const double b[3], double xc, double y1, double y2,
                           double zc, double r) {
    const double D[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double t0 = 0.0, t1 = 1.0;
    if (D[1] == 0.0) {
        if (a[1] < y1 || a[1] > y2) return 0.0;
    } else {
        double u = (y1 - a[1]) / D[1], w = (y2 - a[1]) / D[1];
        if (u > w) {
            const double x = u;
            u = w;
            w = x;
        }
        t0 = tmax2(t0, u);
        t1 = tmin2(t1, w);
    }
    const double x0 = a[0] - xc, z0 = a[2] - zc;
    const double A = D[0] * D[0] + D[2] * D[2], B = x0 * D[0] + z0 * D[2],
                 Cq = x0 * x0 + z0 * z0 - r * r;
    if (A == 0.0) {
        if (Cq > 0.0) return 0.0;
    } else {
        const double disc = B * B - A * Cq;
        if (disc <= 0.0) return 0.0;
        const double sq = sqrt(disc);
        t0 = tmax2(t0, (-B - sq) / A);
        t1 = tmin2(t1, (-B + sq) / A);
    }
    return t1 > t0 ? t1 - t0 : 0.0;
}

SMG_FN double seg_cyl_geo(const double a[3], const double b[3], double xc, double y1, double y2,
                          double zc, double r) {
    const double f = seg_cyl_frac(a, b, xc, y1, y2, zc, r);
    if (!(f > 0.0)) return 0.0;
    const double D[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    return f * sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
}

// Elliptical hole o (semi-axes erx, erz along x', z', the (x, z) offsets from
// the axis turned by theta): in the coordinates x'/erx, z'/erz its section is
// the unit circle.  That map is linear, so a segment's parameter range inside
// is the mapped segment's range inside the unit cylinder.
SMG_FN double seg_ell_geo(const double a[3], const double b[3], const Hole<double>& o) {
    const double ax = a[0] - o.xc, az = a[2] - o.zc, bx = b[0] - o.xc, bz = b[2] - o.zc;
    const double am[3] = {(ax * o.ec + az * o.es) / o.erx, a[1], (-ax * o.es + az * o.ec) / o.erz};
    const double bm[3] = {(bx * o.ec + bz * o.es) / o.erx, b[1], (-bx * o.es + bz * o.ec) / o.erz};
    const double f = seg_cyl_frac(am, bm, 0.0, o.y1, o.y2, 0.0, 1.0);
    if (!(f > 0.0)) return 0.0;
    const double D[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    return f * sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
}

// Collimator material along a -> b at E0: each layer's plate chord less the
// chords of its holes (at least 0) with the layer's mu, plus the holes' own
// mu.  att_t: sum of mu_total l; s_pe, s_c: sums with the photoelectric and
// Compton coefficients.
SMG_FN void coll_material(const PeGeom& G, const double a[3], const double b[3], double& att_t,
                          double& s_pe, double& s_c) {
    att_t = s_pe = s_c = 0.0;
    for (int m = 0; m < G.g.n_layers; ++m) {
        const Layer<double>& L = G.layers_d[m];
        const double c = seg_box_geo(a, b, L.lo, L.hi);
        if (!(c > 0.0)) continue;  // a hole of the layer lies inside its plate box
        double holes = 0.0;
        const int first = G.g.ranges[m].first, last = first + G.g.ranges[m].count;
        for (int h = first; h < last; ++h) {
            const Hole<double>& o = G.holes_d[h];
            const double l = o.erx > 0.0 ? seg_ell_geo(a, b, o) : seg_cyl_geo(a, b, o.xc, o.y1, o.y2, o.zc, o.r);
            holes += l;
            att_t += o.mu * l;
            s_pe += o.mupe * l;
            s_c += o.muc * l;
        }
        const double wl = tmax2(c - holes, 0.0);
        att_t += L.mu * wl;
        s_pe += L.mupe * wl;
        s_c += L.muc * wl;
    }
}

// mu(E) tables (an extension with no reference counterpart): per process
// (0 crystal photoelectric, 1 crystal Compton, 2 collimator photoelectric,
// 3 collimator Compton) log mu on increasing log-energy nodes, linear in
// (log E, log mu) between nodes and continuing the end segments beyond them;
// lmu0 holds each table at E0.  n = 0: the reference's power laws.
constexpr int kMaxMuNodes = 256;
struct MuTables {
    int n;
    double le[kMaxMuNodes];
    double lmu[4][kMaxMuNodes];
    double lmu0[4];
};

SMG_FN double mu_table_log(const MuTables& t, int which, double le) {
    int lo = 0, hi = t.n - 1;
    while (hi - lo > 1) {  // the last node not above le, within [0, n - 2]
        const int mid = (lo + hi) >> 1;
        if (t.le[mid] <= le) lo = mid;
        else hi = mid;
    }
    const double f = (le - t.le[lo]) / (t.le[lo + 1] - t.le[lo]);
    return t.lmu[which][lo] + f * (t.lmu[which][lo + 1] - t.lmu[which][lo]);
}

// mu(E')/mu(E0) of the photoelectric (r3) and Compton (r1) processes of the
// crystal: the tables when given, else (E0/E')^3 and E0/E' (scatter.cu:1371-1375).
SMG_FN void mu_ratios(const MuTables* t, double e0, double ep, double& r3, double& r1) {
    if (!t || t->n == 0) {
        r1 = e0 / ep;
        r3 = e0 * e0 * e0 / ep / ep / ep;
        return;
    }
    const double le = log(ep);
    r3 = exp(mu_table_log(*t, 0, le) - t->lmu0[0]);
    r1 = exp(mu_table_log(*t, 1, le) - t->lmu0[1]);
}

// The tables in the form the FP32 term evaluates them.  E' lies in
// [E0 / (1 + 2 alpha), E0], so each process's ratio mu(E')/mu(E0) is a
// piecewise power law in L = ln(E0/E') on the few table segments that range
// meets: on segment j, ln(mu(E')/mu(E0)) = c[p][j] - s[p][j] L, with lb the
// segment breakpoints in L (ascending).  The FP32 term forms L from
// E0/E' = 1 + alpha (1 - cos theta) and needs no FP64 log or exp: its ratios
// are good to ~3e-7 relative (the FP64 term keeps the table search in FP64).
// n = 0: not available (too many segments in range); the FP32 term then
// takes the FP64 form.
constexpr int kMaxMuSegF = 24;
struct MuTablesF {
    int n;                       // segments
    float lb[kMaxMuSegF];        // lb[j], j < n - 1: upper end of segment j in L
    float c[4][kMaxMuSegF];
    float s[4][kMaxMuSegF];
};

// Segment of L (branch-free count of the breakpoints below it).
SMG_FN int mu_seg_f(const MuTablesF& t, float L) {
    int j = 0;
    for (int k = 0; k < t.n - 1; ++k) j += L > t.lb[k];
    return j;
}

SMG_FN float mu_ratio_f(const MuTablesF& t, int which, int j, float L) {
    return expf(t.c[which][j] - t.s[which][j] * L);
}

// Host: the FP32 form of t for photopeak e0 (alpha = e0 / 511).
inline void make_mu_tables_f(const MuTables& t, double e0, MuTablesF& f) {
    f.n = 0;
    if (t.n < 2) return;
    const double alpha = e0 / 511.0, lmax = std::log(1.0 + 2.0 * alpha) * (1.0 + 1e-6), le0 = std::log(e0);
    double br[kMaxMuNodes];
    int nb = 0;
    for (int k = t.n - 1; k >= 0; --k) {  // nodes inside (E0 e^-lmax, E0), ascending in L
        const double L = le0 - t.le[k];
        if (L > 0.0 && L < lmax) br[nb++] = L;
    }
    if (nb + 1 > kMaxMuSegF) return;
    for (int j = 0; j <= nb; ++j) {
        const double a = j == 0 ? 0.0 : br[j - 1], b = j == nb ? lmax : br[j];
        const double le = le0 - 0.5 * (a + b);  // inside segment j
        int lo = 0, hi = t.n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (t.le[mid] <= le) lo = mid;
            else hi = mid;
        }
        if (j < nb) f.lb[j] = float(br[j]);
        for (int p = 0; p < 4; ++p) {
            const double slope = (t.lmu[p][lo + 1] - t.lmu[p][lo]) / (t.le[lo + 1] - t.le[lo]);
            f.c[p][j] = float(t.lmu[p][lo] + slope * (le0 - t.le[lo]) - t.lmu0[p]);
            f.s[p][j] = float(slope);
        }
    }
    f.n = nb + 1;
}

struct ScatterPhys {
    double e0, lo, hi;     // photopeak energy and energy window (keV)
    double norm;           // sum of KN sin over [0, pi] (scatter.cu:1169)
    double res;            // energy resolution of bar d at E0
    double rcell;          // bounding radius of a detection sub-cell
    double mupe, muc;      // bar d
    const MuTables* tab = nullptr;  // mu(E) tables, or power laws
    const MuTablesF* tabf = nullptr;  // their FP32 form (FP32 term), or none
    const double* knt = nullptr;    // KN prefix table (FP32 path), or the node loop
    double c_rej = -2.0;            // cos(theta) below this is outside the window for
                                    // sure (window_reject_cos with the path's margin)
};

// Thresholds with their reference values.
constexpr double kFwhm = (double)2.35482f;  // scatter.cu:1026, :1263
constexpr double kKnStep = (double)0.01f;   // scatter.cu:1338 (0.01 passed as float)
constexpr double kTwoPi = 2.0 * 3.14159265358979323846;

// The lower 2-sigma window test of a term (scatter.cu:1263-1269) rejects it
// when e0 g (1 + res sqrt(g) 2/FWHM) <= lo, g = E'/E0 = 1 / (1 + alpha (1 -
// cos theta)).  The left side increases with g and g with cos theta, so the
// test is cos theta <= c_lo for one c_lo per (E0, window, bar resolution),
// found here by bisection in FP64.  Returns c_lo less `margin`, which covers
// the rounding of the cosine the caller compares (FP32: below 1e-6 for the
// unit vectors of a term, margin 1e-5; FP64: the reference's acos/cos round
// trip, margin 1e-9), so a term below it is rejected by the full test too;
// -2 when nothing is rejected.
SMG_FN double window_reject_cos(const ScatterPhys& ph, double margin) {
    const double alpha = ph.e0 / 511.0, r = ph.res * (2.0 / kFwhm);
    if (!(ph.lo > 0.0) || !(alpha > 0.0)) return -2.0;
    if (ph.e0 * (1.0 + r) <= ph.lo) return 2.0;  // every term is below the window
    double g0 = 0.0, g1 = 1.0;  // e0 g (1 + r sqrt g) <= lo at g0, > lo at g1
    for (int it = 0; it < 80; ++it) {
        const double gm = 0.5 * (g0 + g1);
        if (ph.e0 * gm * (1.0 + r * sqrt(gm)) <= ph.lo) g0 = gm;
        else g1 = gm;
    }
    const double c_lo = 1.0 - (1.0 / g0 - 1.0) / alpha;
    if (c_lo < -1.5) return -2.0;
    return c_lo - margin;
}

// window_reject_cos(ph, margin) from its value at margin 0: the special
// values -2 (nothing rejected) and 2 (everything rejected) carry no margin.
SMG_FN double window_cut_margin(double cut, double margin) {
    return (cut < -1.5 || cut > 1.5) ? cut : cut - margin;
}

// Klein-Nishina in the reference's factorised form (scatter.cu:962-977), from
// cos(theta).
template <class T>
SMG_FN T kn_of_cos(T c, T alpha) {
    T f1 = alpha * (T(1) - c);
    T f2 = T(1) + c * c;
    return f2 / (T(1) + f1) / (T(1) + f1) * (T(1) + f1 * f1 / f2 / (T(1) + f1));
}

// Inclusive rectangle rule of KN sin on lo + i*step, i = 0..n with
// n = (int)((hi - lo)/step) (scatter.cu:980-993).  FP32: flags a step count
// within rounding distance of an integer, and takes cos and sin of the nodes
// from an angle-addition recurrence in FP64 started at the FP32 cos and sin of
// lo (four FMAs a step instead of two transcendental calls; the rotation keeps
// the start's ~1e-7 error and drifts by about one FP64 ulp per step).  FP64:
// cos and sin of every node, as the reference.
template <class T>
SMG_FN T kn_sum(T lo, T hi, T alpha, bool& near) {
    const T step = T(kKnStep);
    T x = (hi - lo) / step;
    int n = (int)x;
    if constexpr (kGuarded<T>) {
        T fr = x - T(n);
        near |= (fr < T(2e-4) && n > 0) | (fr > T(1) - T(2e-4));
        // cos and sin of the step (double)0.01f = 0.009999999776482582
        const double cst = 0.99995000041890041468, snt = 0.0099998331106604225523;
        // start values in FP32 (the sum is FP32): an FP64 cos/sin here costs
        // more registers than the whole FP32 term
        float cf, sf;
        sincosf(float(lo), &sf, &cf);
        double c = cf, sv = sf;
        KahanSum<T> s;
        for (int i = 0; i <= n; ++i) {
            s.add(kn_of_cos(T(c), alpha) * T(sv) * step);
            const double c2 = c * cst - sv * snt;
            sv = sv * cst + c * snt;
            c = c2;
        }
        return s.value();
    } else {
        KahanSum<T> s;
        for (int i = 0; i <= n; ++i) {
            T th = lo + T(i) * step;
            T c = cos(th), sn = sin(th);
            s.add(kn_of_cos(c, alpha) * sn * step);
        }
        return s.value();
    }
}

// KN rectangle sums from a prefix table.  The reference's sum has nodes
// lo + i step, i = 0..n; with lo = phi + k0 step (0 <= phi < step) it is
// P(phi, k0 + n) - P(phi, k0 - 1), P(phi, k) = sum_{i <= k} g(phi + i step),
// g(t) = KN(cos t) sin t step.  The table holds P in FP64 for kKnPhase + 1
// phases phi_j = j step / kKnPhase (rows) and k < kKnSteps nodes (columns);
// the sum is interpolated linearly in phi between rows j and j + 1 (error
// ~(step / kKnPhase)^2 / 8 times the sum's second derivative, ~1e-9 relative).
constexpr int kKnPhase = 256;
constexpr int kKnSteps = 320;  // nodes up to 3.19 rad; polar angles stay below pi

// Row j of the table for alpha = E0 / 511 (FP64, two-sum compensated).
SMG_FN void kn_prefix_row(double alpha, int j, double* row) {
    const double step = kKnStep;
    const double phi = step * (double)j / (double)kKnPhase;
    KahanSum<double> s;
    for (int k = 0; k < kKnSteps; ++k) {
        const double th = phi + (double)k * step;
        s.add(kn_of_cos<double>(cos(th), alpha) * sin(th) * step);
        row[k] = s.value();
    }
}

// The rectangle sum over [lo, hi] from the table; n as the FP32 node loop
// forms it (with the same near-integer flag).  A node range outside the
// table flags the term for FP64.
SMG_FN double kn_sum_table(const double* tab, float lo, float hi, bool& near) {
    const float stepf = float(kKnStep);
    const float x = (hi - lo) / stepf;
    const int n = (int)x;
    const float fr = x - float(n);
    near |= (fr < 2e-4f && n > 0) | (fr > 1.0f - 2e-4f);
    const double u = (double)lo * (1.0 / kKnStep);
    const int k0 = (int)floor(u);
    double ph = (u - (double)k0) * kKnPhase;
    int j = (int)ph;
    if (j >= kKnPhase) j = kKnPhase - 1;
    if (j < 0) j = 0;
    const double w = ph - (double)j;
    const int k1 = k0 + n;
    if (k0 < 0 || k1 >= kKnSteps) {
        near = true;
        return 0.0;
    }
    const double* r0 = tab + (size_t)j * kKnSteps;
    const double* r1 = r0 + kKnSteps;
    const double a0 = r0[k1] - (k0 > 0 ? r0[k0 - 1] : 0.0);
    const double a1 = r1[k1] - (k0 > 0 ? r1[k0 - 1] : 0.0);
    return a0 + w * (a1 - a0);
}

// Angle between u and w (vectors src -> a and a -> det), used for the polar
// range.  FP64 follows the reference, acos of the clamped cosine
// (scatter.cu:996-1013); FP32 uses atan2(|u x w|, u.w), which keeps its
// precision at small angles where acos of a rounded cosine does not.
template <class T>
SMG_FN T cone_angle(const T u[3], const T w[3]) {
    T dot = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    if constexpr (std::is_same<T, double>::value) {
        T nu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        T nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        T c = dot / (nu * nw);
        if (c > T(1)) c = T(1);
        if (c < T(-1)) c = T(-1);
        return acos(c);
    } else {
        T cx = u[1] * w[2] - u[2] * w[1];
        T cy = u[2] * w[0] - u[0] * w[2];
        T cz = u[0] * w[1] - u[1] * w[0];
        return atan2(sqrt(cx * cx + cy * cy + cz * cz), dot);
    }
}

// Per (scatter point, sub-cell) quantities shared by the lanes.
struct ScatterCell {
    double qs[3];        // scatter point, world
    double qss[3];       // scatter point, frame of bar d
    double q[3];         // sub-cell centre, world
    double lo[3], hi[3]; // sub-cell box, frame of bar d; corner c has lo[k] or hi[k]
                         // on axis k as bit k of c is clear or set
    PairTerm pt;
};

// True when the term of voxel p, scatter point sc.qs and sub-cell centre sc.q
// lies below the energy window by the per-row cosine threshold ph.c_rej
// (window_reject_cos): FP32 a.b < c_rej |a| |b|, or the FP64 cosine below
// it.  Such a term is 0 (scatter_term rejects it too); callers use this to
// skip the term, and loading its weight, early.
template <class T>
SMG_FN bool below_window(const ScatterPhys& ph, const double p[3], const ScatterCell& sc) {
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = sc.qs[k] - p[k];
        b[k] = sc.q[k] - sc.qs[k];
    }
    if constexpr (std::is_same<T, float>::value) {
        const float ax = float(a[0]), ay = float(a[1]), az = float(a[2]);
        const float bx = float(b[0]), by = float(b[1]), bz = float(b[2]);
        const float ab = ax * bx + ay * by + az * bz;
        const float nn = (ax * ax + ay * ay + az * az) * (bx * bx + by * by + bz * bz);
        return ab < float(ph.c_rej) * tsqrt(nn);
    } else {
        const double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        const double na2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double nb2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        return dot / (sqrt(na2) * sqrt(nb2)) < ph.c_rej;
    }
}

// One (voxel, scatter point, sub-cell) term.  wgt is the weight of the
// scatter point for this voxel.  Returns 0 for terms outside the 2-sigma
// energy window (scatter.cu:1263-1269).
template <class T>
SMG_FN T scatter_term(const ScatterPhys& ph, const double p[3], const double ps[3],
                      const ScatterCell& sc, T wgt, bool& near) {
    // Cone angle at the scatter point (scatter.cu:1254), scattered energy
    // (scatter.cu:1016-1023) and energy-window probability (scatter.cu:1025-1033).
    // The window probability sits on the tail of a Gaussian in E', which
    // multiplies the relative error of E' by ~50: E' and the window arguments
    // are formed in FP64 on both paths.  The FP64 path follows the reference
    // (acos of the clamped cosine, then cos of the angle); the fast path takes
    // cos and sin of the angle straight from the vectors.
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = sc.qs[k] - p[k];
        b[k] = sc.q[k] - sc.qs[k];
    }
    if constexpr (std::is_same<T, float>::value) {
        // FP32 pre-test of the lower window cut (most terms of a scatter
        // matrix fall below the window): cos theta < c_rej  <=>
        // a.b < c_rej |a| |b|; a term near the cut goes on to the FP64 test
        const float ax = float(a[0]), ay = float(a[1]), az = float(a[2]);
        const float bx = float(b[0]), by = float(b[1]), bz = float(b[2]);
        const float ab = ax * bx + ay * by + az * bz;
        const float nn = (ax * ax + ay * ay + az * az) * (bx * bx + by * by + bz * bz);
        if (ab < float(ph.c_rej) * tsqrt(nn)) return T(0);
        // mu(E) tables without an FP32 form (more segments in range than it
        // holds): the whole term goes to the FP64 path, which keeps FP64
        // logarithms and exponentials out of the FP32 term
        if (ph.tab && ph.tab->n > 0 && !(ph.tabf && ph.tabf->n > 0)) {
            near = true;
            return T(0);
        }
    }
    const double alpha = ph.e0 / 511.0;
    double omc, st;  // 1 - cos(theta), sin(theta)
    {
        double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        double na2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        double nb2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        if constexpr (std::is_same<T, double>::value) {
            double c = dot / (sqrt(na2) * sqrt(nb2));
            if (c < ph.c_rej) return T(0);  // below the window (the test further down)
            if (c > 1.0) c = 1.0;
            if (c < -1.0) c = -1.0;
            double theta = acos(c);
            omc = 1.0 - cos(theta);
            st = sin(theta);
        } else {
            double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
                   cz = a[0] * b[1] - a[1] * b[0];
            double inv = 1.0 / sqrt(na2 * nb2);
            double c = dot * inv;
            double s2 = (cx * cx + cy * cy + cz * cz) * inv * inv;
            st = sqrt(s2);
            omc = c < 0.5 ? 1.0 - c : s2 / (1.0 + c);
        }
    }
    // E' = E0 g with g = 1 / (1 + alpha (1 - cos)); the FP32 path forms the
    // window from g and one reciprocal, the FP64 path as the reference does
    double ep, resp, z1, z2;
    if constexpr (std::is_same<T, double>::value) {
        ep = ph.e0 / (1.0 + alpha * omc);
        resp = ph.res * sqrt(ep / ph.e0);
        if (ep * (1 + 2 * resp / kFwhm) <= ph.lo || ep * (1 - 2 * resp / kFwhm) >= ph.hi)
            return T(0);  // 2-sigma energy-window early out (scatter.cu:1263-1269)
        double sigma = resp / kFwhm * ep;
        double s2 = sigma * sqrt(2.0);
        z1 = (ph.lo - ep) / s2;
        z2 = (ph.hi - ep) / s2;
    } else {
        // sigma sqrt(2) = resp E' sqrt(2) / FWHM
        constexpr double k2 = 2.0 / kFwhm, ks2 = 1.4142135623730950488 / kFwhm;
        const double g = 1.0 / (1.0 + alpha * omc);
        ep = ph.e0 * g;
        resp = ph.res * sqrt(g);
        if (ep * (1 + resp * k2) <= ph.lo || ep * (1 - resp * k2) >= ph.hi)
            return T(0);  // 2-sigma energy-window early out (scatter.cu:1263-1269)
        const double inv_s2 = 1.0 / (resp * ks2 * ep);  // 1 / (sigma sqrt 2)
        z1 = (ph.lo - ep) * inv_s2;
        z2 = (ph.hi - ep) * inv_s2;
    }
    T pe_win;
    if constexpr (std::is_same<T, double>::value) {
        pe_win = 0.5 * (erf(z2) - erf(z1));
    } else {  // erfc keeps the difference accurate when both z lie on one side
        float f1 = float(z1), f2 = float(z2);
        if (z1 >= 0.0)
            pe_win = 0.5f * (erfcf(f1) - erfcf(f2));
        else if (z2 <= 0.0)
            pe_win = 0.5f * (erfcf(-f2) - erfcf(-f1));
        else
            pe_win = 0.5f * (erff(f2) - erff(f1));
    }

    // azimuthal range from the sub-cell's bounding sphere (scatter.cu:1275-1287);
    // asin is ill-conditioned next to 1, where 1 - r is formed in FP64
    T phi;
    {
        const double L = sc.pt.len, rdet = ph.rcell;
        if (L * st * 2.0 <= rdet) {
            phi = T(kTwoPi);
        } else {
            double r;
            if constexpr (std::is_same<T, double>::value)
                r = rdet / 2.0 / L / st;
            else
                r = rdet / (2.0 * L * st);
            if constexpr (std::is_same<T, double>::value) {
                phi = 4.0 * asin(r <= 1.0 ? r : 1.0);
            } else if (r < 0.75) {
                phi = 4.0f * asinf(float(r));
            } else {  // asin(r) = pi/2 - 2 asin(sqrt((1 - r)/2))
                phi = float(kTwoPi) - 8.0f * asinf(float(sqrt((1.0 - r) * 0.5)));
            }
        }
    }
    const T alpha_t = T(alpha);

    // polar range: cone angles to the 8 sub-cell corners (scatter.cu:1290-1335)
    T us[3];
    for (int k = 0; k < 3; ++k) us[k] = T(sc.qss[k] - ps[k]);
    T tmin = T(0), tmax = T(0);
    if constexpr (std::is_same<T, double>::value) {
        for (int c = 0; c < 8; ++c) {
            T wc[3];
            for (int k = 0; k < 3; ++k) wc[k] = ((c >> k) & 1 ? sc.hi[k] : sc.lo[k]) - sc.qss[k];
            T th = cone_angle<T>(us, wc);
            if (c == 0 || th > tmax) tmax = th;
            if (c == 0 || th < tmin) tmin = th;
        }
    } else {
        // The smallest and largest angle by exact ordering, then two atan2: for
        // angles a, b in [0, pi] given as (|u x w|, u.w) = r (sin, cos),
        // a < b exactly when sin(b - a) > 0, i.e. s_b c_a - c_b s_a > 0.
        float s_lo = 0.0f, c_lo = 1.0f, s_hi = 0.0f, c_hi = 1.0f;
        float wlo[3], whi[3];  // corner offsets from the scatter point
        for (int k = 0; k < 3; ++k) {
            wlo[k] = float(sc.lo[k] - sc.qss[k]);
            whi[k] = float(sc.hi[k] - sc.qss[k]);
        }
        for (int c = 0; c < 8; ++c) {
            float w[3];
            for (int k = 0; k < 3; ++k) w[k] = (c >> k) & 1 ? whi[k] : wlo[k];
            const float cx = us[1] * w[2] - us[2] * w[1], cy = us[2] * w[0] - us[0] * w[2],
                        cz = us[0] * w[1] - us[1] * w[0];
            // (r sin, r cos) with r = |u| |w| > 0: the sign tests and atan2 do
            // not depend on r
            const float sn = tsqrt(cx * cx + cy * cy + cz * cz);
            const float cn = us[0] * w[0] + us[1] * w[1] + us[2] * w[2];
            if (c == 0 || sn * c_hi - cn * s_hi > 0.0f) {
                s_hi = sn;
                c_hi = cn;
            }
            if (c == 0 || s_lo * cn - c_lo * sn > 0.0f) {
                s_lo = sn;
                c_lo = cn;
            }
        }
        tmin = atan2f(s_lo, c_lo);
        tmax = atan2f(s_hi, c_hi);
    }
    // the line voxel -> scatter point, extended 1000 mm, through the sub-cell.
    // FP32 path: only its test against 0.001 is used, and a line that passes
    // the sub-cell's centre farther than its bounding radius + 0.01 mm has
    // chord 0 in the reference too (its small-component branches move the
    // line by less than 0.001 mm), so the chord is formed only otherwise.
    T lthru = T(0);
    bool lthru_needed = true;
    if constexpr (std::is_same<T, float>::value) {
        float w[3];
        for (int k = 0; k < 3; ++k) w[k] = float(0.5 * (sc.lo[k] + sc.hi[k]) - sc.qss[k]);
        const float cx = w[1] * us[2] - w[2] * us[1], cy = w[2] * us[0] - w[0] * us[2],
                    cz = w[0] * us[1] - w[1] * us[0];
        const float u2 = us[0] * us[0] + us[1] * us[1] + us[2] * us[2];
        const float R = float(ph.rcell) + 0.01f;
        lthru_needed = !(cx * cx + cy * cy + cz * cz > R * R * u2);
    }
    if (lthru_needed) {
        double dis2 = 0.0;
        double D[3];
        for (int k = 0; k < 3; ++k) D[k] = sc.qss[k] - ps[k];
        dis2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
        double dis = sqrt(dis2);
        if constexpr (std::is_same<T, float>::value) {
            // De = D (1 + 1000 / |D|), so |De| - |D| = 1000 up to rounding,
            // which the FP32 end parameter does not see
            const double f = 1.0 + 1000.0 / dis;
            const float De[3] = {float(D[0] * f), float(D[1] * f), float(D[2] * f)};
            SegDir<float> se;
            float tde = 4.0f * kUnit<float> * (tabs(De[0]) + tabs(De[1]) + tabs(De[2]));
            segdir_init<float>(se, De[0], De[1], De[2], tde, near);
            float of[3] = {float(sc.qss[0]), float(sc.qss[1]), float(sc.qss[2])};
            float lo[3] = {float(sc.lo[0]), float(sc.lo[1]), float(sc.lo[2])};
            float hi[3] = {float(sc.hi[0]), float(sc.hi[1]), float(sc.hi[2])};
            const float af[3] = {float(ps[0]), float(ps[1]), float(ps[2])};
            lthru = seg_box_local<true>(se, of, float(double(kSlabEps<float>) - dis), 1000.0f, lo, hi,
                                        near, af);
        } else {
            double De[3];
            for (int k = 0; k < 3; ++k) De[k] = (sc.qss[k] + 1000.0 * D[k] / dis) - ps[k];
            SegDir<double> se;
            segdir_init<double>(se, De[0], De[1], De[2], 0.0, near);
            lthru = seg_box<double>(se, ps, sc.lo, sc.hi, near);
        }
    }
    if (g_gt(lthru, T(0.001f), T(1e-5), near)) tmin = T(0);
    T K;
    if constexpr (std::is_same<T, double>::value)
        K = kn_sum<T>(tmin, tmax, alpha_t, near) / T(ph.norm);
    else if (ph.knt)
        K = float(kn_sum_table(ph.knt, tmin, tmax, near) / ph.norm);
    else
        K = kn_sum<T>(tmin, tmax, alpha_t, near) * float(1.0 / ph.norm);

    // second leg and absorption at E' (scatter.cu:1371-1427); the FP32 path
    // takes E0/E' = 1 + alpha (1 - cos) for the power laws
    // (FP32: tables here always have their FP32 form, see the entry test)
    double e3, e1;
    const bool tables = ph.tab && ph.tab->n > 0;
    float Lf = 0.0f;
    int jf = 0;
    if constexpr (std::is_same<T, float>::value) {
        if (!tables) {
            e1 = 1.0 + alpha * omc;
            e3 = e1 * e1 * e1;
        } else {
            Lf = logf(float(1.0 + alpha * omc));  // ln(E0/E')
            jf = mu_seg_f(*ph.tabf, Lf);
            e3 = mu_ratio_f(*ph.tabf, 0, jf, Lf);
            e1 = mu_ratio_f(*ph.tabf, 1, jf, Lf);
        }
    } else {
        mu_ratios(ph.tab, ph.e0, ep, e3, e1);
    }
    double att2 = e3 * sc.pt.s_pe + e1 * sc.pt.s_c;
    if (sc.pt.s_wpe != 0.0 || sc.pt.s_wc != 0.0) {  // collimator material (grid scatterers)
        double w3 = e3, w1 = e1;
        if (tables) {
            if constexpr (std::is_same<T, float>::value) {
                w3 = mu_ratio_f(*ph.tabf, 2, jf, Lf);
                w1 = mu_ratio_f(*ph.tabf, 3, jf, Lf);
            } else {
                const double le = log(ep);
                w3 = exp(mu_table_log(*ph.tab, 2, le) - ph.tab->lmu0[2]);
                w1 = exp(mu_table_log(*ph.tab, 3, le) - ph.tab->lmu0[3]);
            }
        }
        att2 += w3 * sc.pt.s_wpe + w1 * sc.pt.s_wc;
    }
    double mupe = ph.mupe * e3;
    double mut = mupe + ph.muc * e1;
    T ea = exp_neg<T>(sc.pt.la * mut);
    if constexpr (std::is_same<T, double>::value) {
        T absf = (T(1) - ea) * T(mupe / mut);
        return wgt * phi / T(kTwoPi) * K * pe_win * exp_neg<T>(att2) * absf;
    } else {
        T absf = (T(1) - ea) * tdiv(float(mupe), float(mut));
        return wgt * phi * float(1.0 / kTwoPi) * K * pe_win * exp_neg<T>(att2) * absf;
    }
}

}  // namespace smg
