// Bit predicates and bit-order walk of the geometry bitmaps, for one lane.
//
// Crystal bitmap (ScatterGen_RayTracing_CircularHole/scatter.cu:1440-1540):
// bit (i<j, k) says bar k lies between bars i and j by an angular
// bounding-sphere test seen from i or from j; bits of the unordered pair
// {i, j} start at (i N - i(i+1)/2 + j) N (scatter.cu:1040-1059).
// Collimator bitmap (scatter.cu:1939-2009): bit (h, j, k) says bar k lies on
// the ray from layer-0 hole h toward bar j; bit index h N^2 + j N + k.
//
// A row is the N bits of one pair (i, j) or one (hole h, bar j); along a row
// only k changes.  The FP32 test keeps the row's terms (the unit vector from i
// to j and R_j / L_ij) and forms per bit k - i in FP64 (rounded once to FP32,
// so the direction does not inherit the rounding of ~100 mm coordinates), one
// reciprocal square root and squared comparisons.  A decision within its
// rounding bound is redone with the reference formula in FP64, so every bit
// equals an FP64 evaluation of the reference test.
#pragma once

#include "smg_geometry.hpp"
#include "smg_rays.hpp"

namespace smg {

template <class T>
SMG_FN T bm_dist(T ax, T ay, T az, T bx, T by, T bz) {
    T dx = ax - bx, dy = ay - by, dz = az - bz;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// |u_{k<-i} - u_{j<-i}| <= R_k / L_ik + R_j / L_ij, the test seen from i
// (scatter.cu:1518-1533), as the reference writes it.
template <class T>
SMG_FN bool bm_seen_from(const Sphere<T>& i, const Sphere<T>& j, const Sphere<T>& k, bool& near) {
    T lij = bm_dist(i.x, i.y, i.z, j.x, j.y, j.z);
    T lik = bm_dist(k.x, k.y, k.z, i.x, i.y, i.z);
    T ux = (k.x - i.x) / lik - (j.x - i.x) / lij;
    T uy = (k.y - i.y) / lik - (j.y - i.y) / lij;
    T uz = (k.z - i.z) / lik - (j.z - i.z) / lij;
    T s = sqrt(ux * ux + uy * uy + uz * uz);
    T lim = k.r / lik + j.r / lij;
    return g_le(s, lim, T(64) * kUnit<T> * (s + lim + T(1)), near);
}

// Crystal predicate: from i, else from j.
template <class T>
SMG_FN bool bm_crystal_pred(const Sphere<T>& i, const Sphere<T>& j, const Sphere<T>& k, bool& near) {
    bool n1 = false;
    bool a = bm_seen_from(i, j, k, n1);
    near |= n1;
    if (a && !n1) return true;
    bool n2 = false;
    bool b = bm_seen_from(j, i, k, n2);
    near |= n2;
    return a || b;
}

// 1 / sqrt(x): the hardware reciprocal square root on the device (1 ulp),
// the operators on the host.
SMG_FN float trsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// The test seen from o in FP32 with the row's u_oj = unit(j - o) and
// c = R_j / L_oj.  Returns 1 or 0, or -1 when s = |unit(k - o) - u_oj| lies
// within 64 u (3 + lim) of lim: the rounding of s is below ~30 u (s <= 2),
// that of lim below ~10 u lim (u = 2^-24), so outside that band the FP32
// decision is the FP64 one.  A NaN (coincident centres) also returns -1.
SMG_FN int bm_seen_from_f(const Sphere<double>& k, double ox, double oy, double oz, float ux0,
                          float uy0, float uz0, float c) {
    const float dx = float(k.x - ox), dy = float(k.y - oy), dz = float(k.z - oz);
    const float rl = trsq(dx * dx + dy * dy + dz * dz);  // 1 / L_ok
    const float ux = dx * rl - ux0, uy = dy * rl - uy0, uz = dz * rl - uz0;
    const float s2 = ux * ux + uy * uy + uz * uz;
    const float lim = float(k.r) * rl + c;
    const float tol = 64.0f * kUnit<float> * (3.0f + lim);
    const float hi = lim + tol, lo = lim - tol;
    if (s2 > hi * hi) return 0;
    if (lo > 0.0f && s2 < lo * lo) return 1;
    return -1;
}

// The row a lane is walking and its FP32 terms.
struct BitRow {
    int64_t a;           // crystal: bar i; collimator: hole h
    int64_t j;           // bar j
    bool live;           // crystal: i < j (rows with i = j hold no set bit)
    double ox, oy, oz;   // centre of i (or h)
    double jx, jy, jz;   // centre of j
    float ux, uy, uz;    // unit(j - o)
    float ci, cj;        // R_o / L_oj, R_j / L_oj
};

// First bit of pair row i: S(i) N with S(i) = i N - i(i-1)/2.
SMG_FN int64_t bm_row_start(int64_t i, int64_t n) { return i * n - i * (i - 1) / 2; }

SMG_FN int64_t bm_total_bits(bool coll, int64_t n, int64_t n_holes) {
    return coll ? n_holes * n * n : n * (n + 1) / 2 * n;
}

// Row and k of bit b (b below the total).
SMG_FN void bm_row_of_bit(bool coll, int64_t b, int64_t n, BitRow& r, int& k) {
    const int64_t row = b / n;
    k = (int)(b - row * n);
    if (coll) {
        r.a = row / n;
        r.j = row - r.a * n;
        return;
    }
    // invert row = S(i) + (j - i)
    const double disc = (2.0 * n + 1) * (2.0 * n + 1) - 8.0 * (double)row;
    int64_t i = (int64_t)floor(((2.0 * n + 1) - sqrt(disc > 0 ? disc : 0.0)) / 2.0);
    if (i < 0) i = 0;
    if (i > n - 1) i = n - 1;
    while (i > 0 && bm_row_start(i, n) > row) --i;
    while (i + 1 < n && bm_row_start(i + 1, n) <= row) ++i;
    r.a = i;
    r.j = row - bm_row_start(i, n) + i;
}

SMG_FN void bm_next_row(bool coll, int64_t n, BitRow& r) {
    if (++r.j == n) {
        ++r.a;
        r.j = coll ? 0 : r.a;
    }
}

SMG_FN void bm_row_setup(bool coll, const Sphere<double>* bars, const Sphere<double>* holes,
                         BitRow& r) {
    const Sphere<double> o = coll ? holes[r.a] : bars[r.a];
    const Sphere<double> J = bars[r.j];
    r.live = coll || r.a < r.j;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.jx = J.x;
    r.jy = J.y;
    r.jz = J.z;
    const float dx = float(J.x - o.x), dy = float(J.y - o.y), dz = float(J.z - o.z);
    const float rl = trsq(dx * dx + dy * dy + dz * dz);
    r.ux = dx * rl;
    r.uy = dy * rl;
    r.uz = dz * rl;
    r.cj = float(J.r) * rl;
    r.ci = float(o.r) * rl;
}

// Bit k of row r.
SMG_FN bool bm_bit(bool coll, const BitRow& r, const Sphere<double>* bars,
                   const Sphere<double>* holes, int k) {
    if (coll) {
        const Sphere<double> K = bars[k];
        const int a = bm_seen_from_f(K, r.ox, r.oy, r.oz, r.ux, r.uy, r.uz, r.cj);
        if (a >= 0) return a != 0;
        bool nn = false;
        return bm_seen_from<double>(holes[r.a], bars[r.j], K, nn);
    }
    if (!r.live || k == r.a || k == r.j) return false;
    const Sphere<double> K = bars[k];
    const int a = bm_seen_from_f(K, r.ox, r.oy, r.oz, r.ux, r.uy, r.uz, r.cj);
    const int b = bm_seen_from_f(K, r.jx, r.jy, r.jz, -r.ux, -r.uy, -r.uz, r.ci);
    if (a == 1 || b == 1) return true;
    if (a == 0 && b == 0) return false;
    bool nn = false;
    return bm_crystal_pred<double>(bars[r.a], bars[r.j], K, nn);
}

}  // namespace smg
