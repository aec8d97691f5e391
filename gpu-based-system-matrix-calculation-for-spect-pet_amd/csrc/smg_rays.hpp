// Ray primitives of the system-matrix kernels, for one lane.
//
// Semantics follow the reference helpers exactly (PEGen_RayTracing_CircularHole/
// PESysMatGen.cu:20-799): the same "direction component below 0.001 counts as
// zero" branches, the same open/closed interval tests and the same thresholds.
// The arithmetic is organised differently:
//   * a segment's direction data (which components count as zero, its length
//     over the active axes and len/D_k per axis) is computed once per ray and
//     shared by every axis-aligned box the ray is tested against;
//   * every threshold decision made in FP32 is checked against a tolerance
//     derived from the rounding error of its operands; a decision inside its
//     tolerance sets `near`, and the caller recomputes the whole ray in FP64.
//     In FP64 the checks compile away.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SMG_FN __host__ __device__ __forceinline__
#else
// Host build (unit tests of the lane arithmetic with g++).
#include <cmath>
#include <cstdint>
#define SMG_FN inline
using std::sqrt;
using std::fabs;
using std::fmax;
using std::fmin;
#endif

#include <type_traits>

#include "smg_geometry.hpp"

namespace smg {

template <class T>
constexpr bool kGuarded = std::is_same<T, float>::value;

// Unit roundoff of T (2^-24 for float).
template <class T>
constexpr T kUnit = std::is_same<T, float>::value ? T(5.9604645e-08) : T(1.1102230246251565e-16);

template <class T>
SMG_FN T tabs(T x) {
    if constexpr (std::is_floating_point<T>::value) return fabs(x);  // an operand modifier
    else return x < T(0) ? -x : x;
}
template <class T>
SMG_FN T tmax2(T a, T b) {
    if constexpr (std::is_floating_point<T>::value) return fmax(a, b);
    else return a >= b ? a : b;
}
template <class T>
SMG_FN T tmin2(T a, T b) {
    if constexpr (std::is_floating_point<T>::value) return fmin(a, b);
    else return a <= b ? a : b;
}

// Quotient and square root.  FP32 on the device: the hardware reciprocal and
// square root (about 1 ulp each, no denormal rescaling); every FP32 decision
// fed by them carries a tolerance of at least 4 ulp.  FP64, and the host
// build of the lane code: the operators.
template <class T>
SMG_FN T tdiv(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<T, float>::value) return a * __builtin_amdgcn_rcpf(b);
#endif
    return a / b;
}
template <class T>
SMG_FN T trcp(T b) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_rcpf(b);
#endif
    return T(1) / b;
}
template <class T>
SMG_FN T tsqrt(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_sqrtf(x);
#endif
    return sqrt(x);
}

// A wave-uniform value held in scalar registers: values every lane holds
// alike (a tile's bounding ball, a sub-cell's context, a row's physics) would
// otherwise take vector registers; the VALU reads them as scalar operands.
// The caller guarantees the value is the same in every active lane.  Host
// build: the value itself.
SMG_FN double uniform_d(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
#else
    return x;
#endif
}
SMG_FN float uniform_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
#else
    return x;
#endif
}

// FP64 reciprocal, and reciprocal and square root, for the FP32 path's FP64
// set-up: the hardware approximations (v_rcp_f64, v_rsq_f64) refined by two
// Newton steps, within a few ulp of the correctly rounded values at a third of
// the instructions of IEEE division and square root.  Arguments: positive
// normal numbers (callers guard the rest).  The host build (lane tests) and
// the all-FP64 path use the operators.
SMG_FN double rcp_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}
SMG_FN double rsq_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    return y * fma(-h * y, y, 1.5);
#else
    return 1.0 / sqrt(x);
#endif
}

// Element k of a 3-vector by selects: indexing with a run-time k would put
// the whole array in scratch memory.
template <class T>
SMG_FN T at3(const T v[3], int k) {
    return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

// Comparisons that flag a decision taken within `tol` of its threshold.
template <class T>
SMG_FN bool g_lt(T x, T y, T tol, bool& near) {
    if constexpr (kGuarded<T>) near |= tabs(x - y) <= tol;
    return x < y;
}
template <class T>
SMG_FN bool g_le(T x, T y, T tol, bool& near) {
    if constexpr (kGuarded<T>) near |= tabs(x - y) <= tol;
    return x <= y;
}
template <class T>
SMG_FN bool g_gt(T x, T y, T tol, bool& near) {
    return !g_le(x, y, tol, near);
}
template <class T>
SMG_FN bool g_ge(T x, T y, T tol, bool& near) {
    return !g_lt(x, y, tol, near);
}

// Reference thresholds, with the value they have in the reference (float
// literals stay float-valued).
template <class T>
constexpr T kSlabEps = T(0.001f);   // PESysMatGen.cu:26, :326
constexpr double kInsideDy = 0.001;  // PESysMatGen.cu:321 (double literal)
constexpr double kCylDy = 0.001;     // PESysMatGen.cu:695
constexpr double kCylT = 0.0001;     // PESysMatGen.cu:732
constexpr double kCylDisc = 0.00001; // PESysMatGen.cu:745

// Direction data of the segment a -> a + D, shared by all axis-aligned boxes
// of one frame.
template <class T>
struct SegDir {
    T D[3];
    T len;      // length over the active axes
    T inv[3];   // len / D_k on the active axes
    T tol_d;    // tolerance of the |D_k| < eps decisions
    uint8_t small[3];
    int8_t ns;  // number of components below eps
};

template <class T>
SMG_FN void segdir_init(SegDir<T>& s, T dx, T dy, T dz, T tol_d, bool& near) {
    const T eps = kSlabEps<T>;
    s.D[0] = dx;
    s.D[1] = dy;
    s.D[2] = dz;
    s.tol_d = tol_d;
    int ns = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.small[k] = g_lt(tabs(s.D[k]), eps, tol_d, near);
        ns += s.small[k];
    }
    s.ns = ns;
    T l2 = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (!s.small[k] || ns == 3) l2 += s.D[k] * s.D[k];
    // With one small component the reference works with the 2-D length of
    // the other two (PESysMatGen.cu:75, :131, :186); with two or three small
    // components the length is not used.
    s.len = tsqrt(l2);
#pragma unroll
    for (int k = 0; k < 3; ++k) s.inv[k] = s.small[k] ? T(0) : tdiv(s.len, s.D[k]);
}

// Slab intersection over the active axes.  Also returns error bounds of tmin
// and tmax: the rounding of (face - a) is amplified by |len / D_k|.
template <class T>
SMG_FN void slab(const SegDir<T>& s, const T a[3], const T lo[3], const T hi[3],
                                     T& tmin, T& tmax, T& emin, T& emax) {
    tmin = -T(INFINITY);
    tmax = T(INFINITY);
    emin = T(0);
    emax = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (s.small[k]) continue;
        T inv = s.inv[k];
        T t_lo = inv < T(0) ? (hi[k] - a[k]) * inv : (lo[k] - a[k]) * inv;
        T t_hi = inv < T(0) ? (lo[k] - a[k]) * inv : (hi[k] - a[k]) * inv;
        T e = (tabs(a[k]) + tmax2(tabs(lo[k]), tabs(hi[k]))) * tabs(inv);
        if (t_lo > tmin) {
            tmin = t_lo;
            emin = e;
        }
        if (t_hi < tmax) {
            tmax = t_hi;
            emax = e;
        }
    }
}

// Tolerance of a slab parameter with error scale e, and the "grazing" test:
// a bound far above the segment length means the entry or exit parameter is
// ill-conditioned in FP32 (near-grazing ray), so the ray is redone in FP64.
template <class T>
SMG_FN T slab_tol(T e) {
    return T(16) * kUnit<T> * e + T(1e-6);
}

// Chord of the segment a -> a + D through the box [lo, hi]; reference
// length_box_ray (PESysMatGen.cu:20-312).
template <class T>
SMG_FN T seg_box(const SegDir<T>& s, const T a[3], const T lo[3], const T hi[3], bool& near) {
    const T eps = kSlabEps<T>;
    if (s.ns == 3) return T(0);
    if (s.ns == 2) {
        // moving axis k; the start must lie in the box's i/j range and the
        // segment must run past both k faces (PESysMatGen.cu:35-66)
        int k = !s.small[0] ? 0 : (!s.small[1] ? 1 : 2);
        int i = k == 0 ? 1 : 0;
        int j = k == 2 ? 1 : 2;
        const T ai = at3(a, i), aj = at3(a, j), ak = at3(a, k);
        const T loi = at3(lo, i), loj = at3(lo, j), lok = at3(lo, k);
        const T hii = at3(hi, i), hij = at3(hi, j), hik = at3(hi, k);
        T b_k = ak + at3(s.D, k);
        T ti = T(4) * kUnit<T> * (tabs(ai) + tabs(loi) + tabs(hii)) + T(1e-7);
        T tj = T(4) * kUnit<T> * (tabs(aj) + tabs(loj) + tabs(hij)) + T(1e-7);
        T tk = T(4) * kUnit<T> * (tabs(ak) + tabs(b_k) + tabs(lok) + tabs(hik)) + T(1e-7);
        bool in = g_ge(ai, loi, ti, near) && g_le(ai, hii, ti, near) && g_ge(aj, loj, tj, near) &&
                  g_le(aj, hij, tj, near);
        if (!in) return T(0);
        bool span = (g_le(ak, lok, tk, near) && g_ge(b_k, hik, tk, near)) ||
                    (g_le(b_k, lok, tk, near) && g_ge(ak, hik, tk, near));
        return span ? tabs(hik - lok) : T(0);
    }
    if (s.ns == 1) {
        // the start must lie strictly between the faces of the small axis
        // (PESysMatGen.cu:69-73, :125-128, :180-183)
        int k = s.small[0] ? 0 : (s.small[1] ? 1 : 2);
        const T ak = at3(a, k), lok = at3(lo, k), hik = at3(hi, k);
        T tk = T(4) * kUnit<T> * (tabs(ak) + tabs(lok) + tabs(hik)) + T(1e-7);
        if (g_ge(ak, hik, tk, near) || g_le(ak, lok, tk, near)) return T(0);
    }
    T tmin, tmax, emin, emax;
    slab(s, a, lo, hi, tmin, tmax, emin, emax);
    T chord = tmax - tmin;
    T tol_min = slab_tol(emin), tol_max = slab_tol(emax);
    if (g_lt(chord, eps, tol_min + tol_max, near)) return T(0);
    if constexpr (kGuarded<T>) near |= (emin > T(64) * s.len) | (emax > T(64) * s.len);
    T tl = T(8) * kUnit<T> * s.len;
    if (g_ge(tmin, s.len, tol_min + tl, near) || g_ge(tmax, s.len, tol_max + tl, near)) return T(0);
    if (g_le(tmin, eps, tol_min, near) || g_le(tmax, eps, tol_max, near)) return T(0);
    return chord;
}

// Chord of the segment a -> a + D through the box when the end point may lie
// inside it; reference length_box_ray_inside (PESysMatGen.cu:314-686).
template <class T>
SMG_FN T seg_box_inside(const SegDir<T>& s, const T a[3], const T lo[3], const T hi[3],
                            bool& near) {
    const T eps = kSlabEps<T>;
    if (g_lt(tabs(s.D[1]), T(kInsideDy), s.tol_d, near)) return T(0);
    if (s.ns == 3) return T(0);
    if (s.ns == 2) {
        // Two components below eps (only x with z can remain once |D_y| >=
        // 0.001 up to the eps rounding): the chord runs from the entry face
        // to the end point (PESysMatGen.cu:333-382).
        int k = !s.small[0] ? 0 : (!s.small[1] ? 1 : 2);
        int i = k == 0 ? 1 : 0;
        int j = k == 2 ? 1 : 2;
        const T ai = at3(a, i), aj = at3(a, j), ak = at3(a, k);
        const T loi = at3(lo, i), loj = at3(lo, j), lok = at3(lo, k);
        const T hii = at3(hi, i), hij = at3(hi, j), hik = at3(hi, k);
        T b_k = ak + at3(s.D, k);
        T ti = T(4) * kUnit<T> * (tabs(ai) + tabs(loi) + tabs(hii)) + T(1e-7);
        T tj = T(4) * kUnit<T> * (tabs(aj) + tabs(loj) + tabs(hij)) + T(1e-7);
        bool in = g_ge(ai, loi, ti, near) && g_le(ai, hii, ti, near) && g_ge(aj, loj, tj, near) &&
                  g_le(aj, hij, tj, near);
        if (!in) return T(0);
        if (ak < b_k) return b_k - lok;
        if (ak > b_k) return hik - b_k;
        return T(0);
    }
    if (s.ns == 1) {
        int k = s.small[0] ? 0 : (s.small[1] ? 1 : 2);
        const T ak = at3(a, k), lok = at3(lo, k), hik = at3(hi, k);
        T tk = T(4) * kUnit<T> * (tabs(ak) + tabs(lok) + tabs(hik)) + T(1e-7);
        if (g_ge(ak, hik, tk, near) || g_le(ak, lok, tk, near)) return T(0);
    }
    T tmin, tmax, emin, emax;
    slab(s, a, lo, hi, tmin, tmax, emin, emax);
    T chord = tmax - tmin;
    T tol_min = slab_tol(emin), tol_max = slab_tol(emax);
    if constexpr (kGuarded<T>) near |= (emin > T(64) * s.len) | (emax > T(64) * s.len);
    T tl = T(8) * kUnit<T> * s.len;
    // clipping rules of PESysMatGen.cu:415-434
    if (g_lt(chord, eps, tol_min + tol_max, near)) return T(0);
    if (g_ge(tmin, s.len, tol_min + tl, near)) return T(0);
    if (g_le(tmax, eps, tol_max, near)) return T(0);
    bool max_past_end = g_ge(tmax, s.len, tol_max + tl, near);
    bool min_at_start = g_le(tmin, eps, tol_min, near);
    if (max_past_end && !min_at_start) return s.len - tmin;
    if (min_at_start && g_le(tmax, s.len, tol_max + tl, near)) return tmax;
    if (min_at_start && max_past_end) return s.len;
    return chord;
}

// ---- FP32 forms with a local origin ------------------------------------------
//
// A chord through a box 50 mm down the ray is the difference of two slab
// parameters of about 50 mm; in FP32 that difference keeps only ~3e-6 mm.
// The forms below take the slab parameters relative to a point o on the ray
// close to the box (o = start + unit direction * t0, placed by the caller in
// FP64), so the parameters are box-sized and the chord keeps FP32 relative
// precision.  The reference's parameter tests (t >= len, t <= eps) become
// tests against lim_hi = len - t0 and lim_lo = eps - t0.  Segments with a
// direction component below eps, and slabs bound by an axis with
// |len / D_k| > kGrazeInv, are flagged for the FP64 path.
constexpr float kGrazeInv = 32.0f;

template <bool kSkipSmall = false>
SMG_FN void slab_local(const SegDir<float>& s, const float o[3], const float lo[3],
                       const float hi[3], float& tmin, float& tmax, float& emin, float& emax,
                       float& imin, float& imax) {
    tmin = -INFINITY;
    tmax = INFINITY;
    emin = emax = imin = imax = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if constexpr (kSkipSmall) {
            if (s.small[k]) continue;
        }
        float inv = s.inv[k];
        float t_lo = inv < 0.0f ? (hi[k] - o[k]) * inv : (lo[k] - o[k]) * inv;
        float t_hi = inv < 0.0f ? (lo[k] - o[k]) * inv : (hi[k] - o[k]) * inv;
        float ai = tabs(inv);
        float e = (tabs(o[k]) + tmax2(tabs(lo[k]), tabs(hi[k]))) * ai;
        if (t_lo > tmin) {
            tmin = t_lo;
            emin = e;
            imin = ai;
        }
        if (t_hi < tmax) {
            tmax = t_hi;
            emax = e;
            imax = ai;
        }
    }
}

// length_box_ray semantics (PESysMatGen.cu:236-312) with a local origin.
// kOneSmall: a segment with one direction component below eps takes the
// reference's branch for it (PESysMatGen.cu:69-73): the start point a (same
// frame as lo, hi) must lie strictly between that axis' faces, and the slab
// runs over the other two axes in units of their 2-D length.  Otherwise, and
// with two or three small components, the segment is flagged.
template <bool kOneSmall = false>
SMG_FN float seg_box_local(const SegDir<float>& s, const float o[3], float lim_lo, float lim_hi,
                           const float lo[3], const float hi[3], bool& near,
                           const float* a = nullptr) {
    const float eps = kSlabEps<float>;
    if (s.ns != 0) {
        if (!kOneSmall || s.ns > 1) {
            near = true;
            return 0.0f;
        }
        const int k = s.small[0] ? 0 : (s.small[1] ? 1 : 2);
        const float ak = at3(a, k), lok = at3(lo, k), hik = at3(hi, k);
        const float tk = 4.0f * kUnit<float> * (tabs(ak) + tabs(lok) + tabs(hik)) + 1e-7f;
        if (g_ge(ak, hik, tk, near) || g_le(ak, lok, tk, near)) return 0.0f;
    }
    float tmin, tmax, emin, emax, imin, imax;
    slab_local<kOneSmall>(s, o, lo, hi, tmin, tmax, emin, emax, imin, imax);
    float chord = tmax - tmin;
    float tol_min = slab_tol(emin), tol_max = slab_tol(emax);
    if (g_lt(chord, eps, tol_min + tol_max, near)) return 0.0f;
    near |= (imin > kGrazeInv) | (imax > kGrazeInv);
    float tlh = 8.0f * kUnit<float> * tabs(lim_hi) + 1e-6f;
    float tll = 8.0f * kUnit<float> * tabs(lim_lo) + 1e-6f;
    if (g_ge(tmin, lim_hi, tol_min + tlh, near) || g_ge(tmax, lim_hi, tol_max + tlh, near))
        return 0.0f;
    if (g_le(tmin, lim_lo, tol_min + tll, near) || g_le(tmax, lim_lo, tol_max + tll, near))
        return 0.0f;
    return chord;
}

// length_box_ray_inside semantics (PESysMatGen.cu:415-434) with a local
// origin; t0 is the position of o along the segment.
SMG_FN float seg_box_inside_local(const SegDir<float>& s, const float o[3], float t0,
                                  float lim_lo, float lim_hi, const float lo[3],
                                  const float hi[3], bool& near) {
    const float eps = kSlabEps<float>;
    if (g_lt(tabs(s.D[1]), float(kInsideDy), s.tol_d, near)) return 0.0f;
    if (s.ns != 0) {
        near = true;
        return 0.0f;
    }
    float tmin, tmax, emin, emax, imin, imax;
    slab_local(s, o, lo, hi, tmin, tmax, emin, emax, imin, imax);
    float chord = tmax - tmin;
    float tol_min = slab_tol(emin), tol_max = slab_tol(emax);
    float tlh = 8.0f * kUnit<float> * tabs(lim_hi) + 1e-6f;
    float tll = 8.0f * kUnit<float> * tabs(lim_lo) + 1e-6f;
    if (g_lt(chord, eps, tol_min + tol_max, near)) return 0.0f;
    near |= (imin > kGrazeInv) | (imax > kGrazeInv);
    if (g_ge(tmin, lim_hi, tol_min + tlh, near)) return 0.0f;
    if (g_le(tmax, lim_lo, tol_max + tll, near)) return 0.0f;
    bool max_past_end = g_ge(tmax, lim_hi, tol_max + tlh, near);
    bool min_at_start = g_le(tmin, lim_lo, tol_min + tll, near);
    if (max_past_end && !min_at_start) return lim_hi - tmin;
    if (min_at_start && g_le(tmax, lim_hi, tol_max + tlh, near)) return tmax + t0;
    if (min_at_start && max_past_end) return s.len;
    return chord;
}

// Per-ray part of the cylinder chord: the unit direction, the ratios the
// reference forms for every hole, and the reciprocals of k_y and of
// k_x^2 + k_z^2 that replace its per-hole divisions (a change of at most an
// ulp in FP64).
struct CylRay {
    double kx, ky, kz, sxk, szk, qa, inv_ky, inv_qa;
};

SMG_FN void cyl_ray(const double D[3], double len, CylRay& c) {
    const double il = 1.0 / len;
    c.kx = D[0] * il;
    c.ky = D[1] * il;
    c.kz = D[2] * il;
    c.inv_ky = 1.0 / c.ky;
    c.sxk = c.kx * c.inv_ky;
    c.szk = c.kz * c.inv_ky;
    c.qa = c.kx * c.kx + c.kz * c.kz;
    c.inv_qa = 1.0 / c.qa;
}

// cyl_ray from reciprocals the caller already holds: il = 1 / len and
// inv_ky = len / D_y (= 1 / k_y up to rounding); one division left.
SMG_FN void cyl_ray_from(const double D[3], double il, double inv_ky, CylRay& c) {
    c.kx = D[0] * il;
    c.ky = D[1] * il;
    c.kz = D[2] * il;
    c.inv_ky = inv_ky;
    c.sxk = c.kx * inv_ky;
    c.szk = c.kz * inv_ky;
    c.qa = c.kx * c.kx + c.kz * c.kz;
    c.inv_qa = c.qa > 1e-300 ? rcp_fast(c.qa) : 1.0 / c.qa;
}

// Chord through the finite cylinder of radius r, axis parallel to y at
// (xc, zc), spanning y1..y2; reference length_cylinder_ray
// (PESysMatGen.cu:688-799).  Evaluated in FP64: rays through a hole run
// almost along its axis, where kx^2 + kz^2 -> 0 and the quadratic's roots
// lose every FP32 digit.
template <bool kFastSqrt = false>
SMG_FN double cyl_chord(const double a[3], const CylRay& c, double len, double xc, double y1,
                        double y2, double zc, double r) {
    if (fabs(y1 - y2) < kCylDy) return 0.0;
    const double kx = c.kx, kz = c.kz, sxk = c.sxk, szk = c.szk;
    const double a1 = y1 - a[1], a2 = y2 - a[1];
    double xl = a[0] + sxk * a1 - xc;
    double zl = a[2] + szk * a1 - zc;
    double xr = a[0] + sxk * a2 - xc;
    double zr = a[2] + szk * a2 - zc;
    double t_lo = a1 * c.inv_ky;
    double t_hi = a2 * c.inv_ky;
    double r2 = r * r;
    bool in_l = xl * xl + zl * zl <= r2;
    bool in_r = xr * xr + zr * zr <= r2;
    if (in_l && in_r) {
        if (t_lo <= kCylT || t_hi <= kCylT) return 0.0;
        if (t_lo >= len || t_hi >= len) return 0.0;
        return fabs(t_hi - t_lo);
    }
    double x0 = a[0] - xc, z0 = a[2] - zc;
    const double qa = c.qa;
    double cr = kx * z0 - kz * x0;
    double disc = qa * r2 - cr * cr;
    if (disc <= kCylDisc) return 0.0;
    double sq = kFastSqrt ? disc * rsq_fast(disc) : sqrt(disc);
    double qb = kx * x0 + kz * z0;
    double t1 = (-qb - sq) * c.inv_qa;
    double t2 = (-qb + sq) * c.inv_qa;
    if (!in_l && !in_r) {
        if (t1 >= t_lo && t2 <= t_hi) {
            if (t2 <= kCylT || t1 <= kCylT) return 0.0;
            if (t2 >= len || t1 >= len) return 0.0;
            return t2 - t1;
        }
        return 0.0;
    }
    if (in_l) {
        if (t2 <= kCylT || t_lo <= kCylT) return 0.0;
        if (t2 >= len || t_lo >= len) return 0.0;
        return t2 - t_lo;
    }
    if (t_hi <= kCylT || t1 <= kCylT) return 0.0;
    if (t_hi >= len || t1 >= len) return 0.0;
    return t_hi - t1;
}

// Chord through an elliptical hole o: the finite elliptic cylinder
// ((x' - xc)/erx)^2 + ((z' - zc)/erz)^2 <= 1, y1 <= y <= y2, where (x', z') is
// the point turned by theta about the axis (cos, sin = ec, es); reference
// length_ellipticalcylinder_ray (PESysMatGen.cu:801-947), a routine the
// reference defines but never calls.  The ray is the start point a, the unit
// direction (kx, ky, kz) with inv_ky = 1/ky, and len.  After the turn the
// cases and thresholds are the circular chord's; the quadratic is the
// ellipse's, a = kx'^2 erz^2 + kz'^2 erx^2, b = kx' x0' erz^2 + kz' z0' erx^2,
// with the cut b^2 - a c <= 1e-5 (PESysMatGen.cu:887-893), formed as
// erx^2 erz^2 (a - (kx' z0' - kz' x0')^2), the value of b^2 - a c without its
// cancellation.  FP64, as the circle's chord, in every kernel that has it
// compiled in (the FP32 kernels get a variant with it for elliptical holes).
SMG_FN double ell_chord(const Hole<double>& o, double ax, double ay, double az, double kx_,
                        double kz_, double inv_ky, double len) {
    if (fabs(o.y1 - o.y2) < kCylDy) return 0.0;
    const double dx = ax - o.xc, dz = az - o.zc;
    const double x0 = dx * o.ec + dz * o.es, z0 = -dx * o.es + dz * o.ec;
    const double kx = kx_ * o.ec + kz_ * o.es, kz = -kx_ * o.es + kz_ * o.ec;
    const double sxk = kx * inv_ky, szk = kz * inv_ky;
    const double a1 = o.y1 - ay, a2 = o.y2 - ay;
    const double xl = x0 + sxk * a1, zl = z0 + szk * a1;
    const double xr = x0 + sxk * a2, zr = z0 + szk * a2;
    const double t_lo = a1 * inv_ky, t_hi = a2 * inv_ky;
    const double ax2 = o.erx * o.erx, az2 = o.erz * o.erz;
    const bool in_l = xl * xl / ax2 + zl * zl / az2 <= 1.0;
    const bool in_r = xr * xr / ax2 + zr * zr / az2 <= 1.0;
    if (in_l && in_r) {
        if (t_lo <= kCylT || t_hi <= kCylT) return 0.0;
        if (t_lo >= len || t_hi >= len) return 0.0;
        return fabs(t_hi - t_lo);
    }
    const double qa = kx * kx * az2 + kz * kz * ax2;
    const double cr = kx * z0 - kz * x0;
    const double disc = ax2 * az2 * (qa - cr * cr);
    if (disc <= kCylDisc) return 0.0;
    const double sq = sqrt(disc);
    const double qb = kx * x0 * az2 + kz * z0 * ax2;
    const double t1 = (-qb - sq) / qa;
    const double t2 = (-qb + sq) / qa;
    if (!in_l && !in_r) {
        if (t1 >= t_lo && t2 <= t_hi) {
            if (t2 <= kCylT || t1 <= kCylT) return 0.0;
            if (t2 >= len || t1 >= len) return 0.0;
            return t2 - t1;
        }
        return 0.0;
    }
    if (in_l) {
        if (t2 <= kCylT || t_lo <= kCylT) return 0.0;
        if (t2 >= len || t_lo >= len) return 0.0;
        return t2 - t_lo;
    }
    if (t_hi <= kCylT || t1 <= kCylT) return 0.0;
    if (t_hi >= len || t1 >= len) return 0.0;
    return t_hi - t1;
}

// Compensated (two-sum) accumulator.
template <class T>
struct KahanSum {
    T s = T(0), c = T(0);
    SMG_FN void add(T x) {
        T t = s + x;
        T bp = t - s;
        c += (s - (t - bp)) + (x - bp);
        s = t;
    }
    SMG_FN T value() const { return s + c; }
};

}  // namespace smg
