// Per-lane body of the photoelectric (PE) matrix kernel.
//
// One lane owns one matrix element A[d][v] (detector bar d, voxel v) and
// evaluates the reference element formula (PEGen_RayTracing_CircularHole/
// PESysMatGen.cu:951-1364) for one detector sub-cell at a time:
//
//   A[d][v] = sum over sub-cells of  Omega * exp(-att) * (1 - exp(-mu_t l_cell)) mu_pe / mu_t
//
// att = plate gates + hole chords + plate material + other bars + own bar.
// Positions and ray directions are formed in FP64; the bulk of the chord
// arithmetic runs in T (float on the fast path, double on the exact path).
// Hole chords and the plate-material term always run in FP64.
#pragma once

#include "smg_geometry.hpp"
#include "smg_rays.hpp"

#if defined(__HIPCC__)
#define SMG_HD __host__ __device__
#else
#define SMG_HD
#endif

namespace smg {

struct PeGeom {
    GeomDesc g;
    const Bar<float>* bars_f;
    const Bar<double>* bars_d;
    const Hole<float>* holes_f;
    const Hole<double>* holes_d;
    const Layer<float>* layers_f;
    const Layer<double>* layers_d;
    const GroupBox* bar_groups;
    const GroupBox* hole_groups;
};

// Voxel centre, FOV shift, rotation about z by rot * dphi
// (PESysMatGen.cu:1028-1053).
SMG_FN void voxel_position(const GeomDesc& g, int ix, int iy, int iz, double cphi, double sphi,
                           double p[3]) {
    double x = (ix - g.nx / 2.0 + 0.5) * g.wx + g.sx;
    double y = (iy - g.ny / 2.0 + 0.5) * g.wy + g.sy;
    double z = (iz - g.nz / 2.0 + 0.5) * g.wz + g.sz;
    p[0] = x * cphi - y * sphi;
    p[1] = x * sphi + y * cphi;
    p[2] = z;
}

// World -> bar frame: rotation by -beta about y around the bar centre
// (PESysMatGen.cu:1071-1073).
template <class B>
SMG_FN void to_frame(const B& b, const double u[3], double out[3]) {
    double cb = b.cb, sb = -b.sb;  // cos(-beta), sin(-beta)
    double x = u[0] - b.cx, z = u[2] - b.cz;
    out[0] = x * cb - z * sb;
    out[1] = u[1] - b.cy;
    out[2] = x * sb + z * cb;
}

// Bar frame -> world (PESysMatGen.cu:1211-1217).
template <class B>
SMG_FN void from_frame(const B& b, const double s[3], double out[3]) {
    out[0] = b.cx + s[0] * b.cb - s[2] * b.sb;
    out[1] = b.cy + s[1];
    out[2] = b.cz + s[0] * b.sb + s[2] * b.cb;
}

// Sub-cell (a, b, c) of an (nsx, nsy, nsz) split: centre and box in the bar
// frame (PESysMatGen.cu:1207-1209, :1329-1336).
SMG_FN void sub_cell(const Bar<double>& b, int a, int bb, int c, int nsx, int nsy, int nsz,
                     double ctr[3], double lo[3], double hi[3]) {
    ctr[0] = -b.w / 2.0 + (a + 0.5) / nsx * b.w;
    ctr[1] = -b.t / 2.0 + (bb + 0.5) / nsy * b.t;
    ctr[2] = -b.h / 2.0 + (c + 0.5) / nsz * b.h;
    lo[0] = ((double)a / nsx - 0.5) * b.w;
    hi[0] = (((double)a + 1.0) / nsx - 0.5) * b.w;
    lo[1] = ((double)bb / nsy - 0.5) * b.t;
    hi[1] = (((double)bb + 1.0) / nsy - 0.5) * b.t;
    lo[2] = ((double)c / nsz - 0.5) * b.h;
    hi[2] = (((double)c + 1.0) / nsz - 0.5) * b.h;
}

// ---- culling (PESysMatGen.cu:1097-1186) ------------------------------------------
//
// The reference keeps bar k (k != d) for element (d, v) when
//   |u_k - u_d| <= rdiag_k / (2 D_k) + rdiag_d / (2 D_d),
// with u the unit vector from the voxel to a bar centre and D its distance,
// and hole h when |u_h - u_d| <= crit_h / D_h + rdiag_d / (2 D_d).  Neither
// test is conservative for the chord it gates, so the GPU must take exactly
// the same decisions: FP32 with a tolerance, FP64 when inside it.

struct SelfCull {
    double ud[3];
    double crit;
};

SMG_FN void self_cull(const Bar<double>& bd, const double p[3], SelfCull& s) {
    double dx = bd.cx - p[0], dy = bd.cy - p[1], dz = bd.cz - p[2];
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    s.crit = bd.rdiag / dist / 2.0;
    s.ud[0] = dx / dist;
    s.ud[1] = dy / dist;
    s.ud[2] = dz / dist;
}

// Angular distance and threshold of one (voxel, sphere) pair in T.
template <class T>
SMG_FN bool cull_test(T px, T py, T pz, T cx, T cy, T cz, T crit_num, T half, const T ud[3],
                      T crit_self, bool& near) {
    T dx = px - cx, dy = py - cy, dz = pz - cz;
    T dk = tsqrt(dx * dx + dy * dy + dz * dz);
    if constexpr (std::is_same<T, double>::value) {  // the reference's operation order
        T crit = half ? crit_num / dk / T(2) : crit_num / dk;
        T ux = (cx - px) / dk, uy = (cy - py) / dk, uz = (cz - pz) / dk;
        T s = sqrt((uz - ud[2]) * (uz - ud[2]) + (uy - ud[1]) * (uy - ud[1]) + (ux - ud[0]) * (ux - ud[0]));
        T lim = crit + crit_self;
        return g_le(s, lim, T(64) * kUnit<T> * (s + lim + T(1)), near);
    }
    const T rk = trcp(dk);
    T crit = half ? crit_num * rk * T(0.5) : crit_num * rk;
    T ux = -dx * rk, uy = -dy * rk, uz = -dz * rk;
    T s = tsqrt((uz - ud[2]) * (uz - ud[2]) + (uy - ud[1]) * (uy - ud[1]) + (ux - ud[0]) * (ux - ud[0]));
    T lim = crit + crit_self;
    return g_le(s, lim, T(64) * kUnit<T> * (s + lim + T(1)), near);
}

// Exact decision for bar k: FP32 first (fast path), FP64 inside the tolerance.
template <bool kExact>
SMG_FN bool cull_bar(const PeGeom& G, int k, const double p[3], const float pf[3],
                     const SelfCull& sc, const float udf[3], float critf) {
    if constexpr (!kExact) {
        const Bar<float>& b = G.bars_f[k];
        bool near = false;
        bool keep = cull_test<float>(pf[0], pf[1], pf[2], b.cx, b.cy, b.cz, b.rdiag, 1.0f, udf,
                                     critf, near);
        if (!near) return keep;
    }
    const Bar<double>& b = G.bars_d[k];
    bool near = false;
    return cull_test<double>(p[0], p[1], p[2], b.cx, b.cy, b.cz, b.rdiag, 1.0, sc.ud, sc.crit, near);
}

template <bool kExact>
SMG_FN bool cull_hole(const PeGeom& G, int h, const double p[3], const float pf[3],
                      const SelfCull& sc, const float udf[3], float critf) {
    if constexpr (!kExact) {
        const Hole<float>& o = G.holes_f[h];
        bool near = false;
        bool keep = cull_test<float>(pf[0], pf[1], pf[2], o.xc, o.ymid, o.zc, o.crit, 0.0f, udf,
                                     critf, near);
        if (!near) return keep;
    }
    const Hole<double>& o = G.holes_d[h];
    bool near = false;
    return cull_test<double>(p[0], p[1], p[2], o.xc, o.ymid, o.zc, o.crit, 0.0, sc.ud, sc.crit, near);
}

// ---- candidate narrowing for a tile of voxels ------------------------------------------

// Bounding ball of the voxel centres of the tile with first voxel (ix0, iy0,
// iz0) and extent tile[] (clipped at the image edge); the rotation about z
// preserves distances.
SMG_FN void tile_ball(const GeomDesc& g, int ix0, int iy0, int iz0, const int tile[3], double cphi,
                      double sphi, double tc[3], double& rho) {
    const int ex = tmin2(tile[0], g.nx - ix0), ey = tmin2(tile[1], g.ny - iy0),
              ez = tmin2(tile[2], g.nz - iz0);
    double x = (ix0 + 0.5 * (ex - 1) - g.nx / 2.0 + 0.5) * g.wx + g.sx;
    double y = (iy0 + 0.5 * (ey - 1) - g.ny / 2.0 + 0.5) * g.wy + g.sy;
    double z = (iz0 + 0.5 * (ez - 1) - g.nz / 2.0 + 0.5) * g.wz + g.sz;
    tc[0] = x * cphi - y * sphi;
    tc[1] = x * sphi + y * cphi;
    tc[2] = z;
    rho = 0.5 * sqrt(((ex - 1) * g.wx) * ((ex - 1) * g.wx) + ((ey - 1) * g.wy) * ((ey - 1) * g.wy) +
                     ((ez - 1) * g.wz) * ((ez - 1) * g.wz)) + 1e-6;
}

// FP32 copy of the ball for box_meets_frustum, radius rounded up.
SMG_FN void tile_ball_f(const double tc[3], double rho, float tcf[3], float& rhof) {
    for (int k = 0; k < 3; ++k) tcf[k] = float(tc[k]);
    rhof = float(rho) * 1.0001f;
}
//
// Can a segment from a point of the box (q, rq) to a point of the ball
// (tc, rho) meet the box of half sizes (hx, hy, hz) centred at c and turned by
// beta about y (cos and sin in cb, sb)?  rq holds the half sizes of an
// axis-aligned box around q in the frame of the tested box.  Such a segment is
//   (q + e1) + s (tc + e2 - q - e1),  s in [0, 1],  |e1_i| <= rq_i,  |e2| <= rho,
// so per axis it lies within r_i(s) = (1 - s) rq_i + s rho of the axis point
// q + s (tc - q): the box grown by r_i(s) must meet the axis segment.  Per axis
// that is two linear inequalities in s; the test intersects their intervals
// with [0, 1].  The margin m covers the FP32 rounding of the inputs and the
// direction components below 0.001 that the chord functions treat as zero (a
// segment then drifts less than 0.001 off the box along that axis).  A false
// answer means every chord of every such segment through the box is zero.
SMG_FN bool box_meets_frustum(const float q[3], const float rq[3], const float tc[3], float rho,
                              float cx, float cy, float cz, float cb, float sb, float hx, float hy,
                              float hz) {
    float qx = q[0] - cx, qy = q[1] - cy, qz = q[2] - cz;
    float tx = tc[0] - cx, ty = tc[1] - cy, tz = tc[2] - cz;
    // into the box frame (rotation by -beta about y, as to_frame)
    float a[3] = {qx * cb + qz * sb, qy, qz * cb - qx * sb};
    float b[3] = {tx * cb + tz * sb, ty, tz * cb - tx * sb};
    float h[3] = {hx, hy, hz};
    float lo = 0.0f, hi = 1.0f;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float u = b[i] - a[i];
        float dr = rho - rq[i];
        float m = 2e-3f + 1e-5f * (tabs(a[i]) + tabs(b[i]) + h[i] + rho + rq[i]);
        float c0 = u - dr, r0 = h[i] + m + rq[i] - a[i];   // a + s u <= h + r(s)
        float c1 = -u - dr, r1 = h[i] + m + rq[i] + a[i];  // -(a + s u) <= h + r(s)
        if (c0 > 0.0f) hi = tmin2(hi, tdiv(r0, c0));
        else if (c0 < 0.0f) lo = tmax2(lo, tdiv(r0, c0));
        else ok &= r0 >= 0.0f;
        if (c1 > 0.0f) hi = tmin2(hi, tdiv(r1, c1));
        else if (c1 < 0.0f) lo = tmax2(lo, tdiv(r1, c1));
        else ok &= r1 >= 0.0f;
    }
    return ok && lo <= hi + 1e-5f;
}


// Segments from any point of bar d to the tile ball, against a box turned by
// beta_k (cos, sin in cbk, sbk): bar d's half sizes become the axis-aligned
// extents of bar d seen in the frame of that box.
SMG_FN bool bar_tile_may_hit(const Bar<float>& bd, const float tc[3], float rho, float cx,
                             float cy, float cz, float cbk, float sbk, float hx, float hy, float hz) {
    const float cd = tabs(bd.cb * cbk + bd.sb * sbk), sd = tabs(bd.sb * cbk - bd.cb * sbk);
    const float hdx = 0.5f * bd.w * 1.0001f, hdy = 0.5f * bd.t * 1.0001f, hdz = 0.5f * bd.h * 1.0001f;
    const float rq[3] = {cd * hdx + sd * hdz, hdy, sd * hdx + cd * hdz};
    const float q[3] = {bd.cx, bd.cy, bd.cz};
    return box_meets_frustum(q, rq, tc, rho, cx, cy, cz, cbk, sbk, hx, hy, hz);
}

// a[i] = v for a run-time i < N through unrolled selects: indexing a local
// array with a run-time value would place it in scratch memory.
template <int N>
SMG_FN void store_at(double* a, int i, double v) {
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k == i) a[k] = v;
}

// ---- exp(-x) for x >= 0 given in FP64, result in T ---------------------------------
// FP32: e^-x = 2^-(x log2 e) split into integer and fraction in FP64, so the
// rounding of a large x (a ray through 5 mm of tungsten has x ~ 20) does not
// turn into a relative error of the result.
template <class T>
SMG_FN T exp_neg(double x) {
    if constexpr (std::is_same<T, double>::value) {
        return exp(-x);
    } else {
        double y = x * 1.4426950408889634;  // log2(e)
        if (y > 200.0) return 0.0f;
        double n = floor(y);
        float f = (float)(y - n);
        return ldexpf(exp2f(-f), -(int)n);
    }
}

// Work item of block b in a grid of n blocks.  The dispatcher deals
// consecutive blocks round-robin over the 8 XCDs of the chip, each with its
// own L2.  Here the items go out in chunks of kXcdChunk consecutive items,
// chunk j to XCD j % 8 (its blocks b = j % 8 + 8 (kXcdChunk (j / 8) + o),
// o < kXcdChunk, in order): the waves that write neighbouring voxel runs of
// an output row (x-adjacent tiles, 16 B each) share an L2, which merges their
// partial-line stores into whole 128 B lines before they go to HBM, while
// every XCD still draws its chunks from all rows (the work per row varies
// with the bar's place in the detector).  The last n % (8 kXcdChunk) blocks
// keep their own index.  A bijection on [0, n).
constexpr int kXcds = 8;
constexpr int kXcdChunk = 8;
SMG_FN int xcd_item(int b, int n) {
    const int span = kXcds * kXcdChunk;
    if (b >= n / span * span) return b;
    const int x = b % kXcds, local = b / kXcds;
    return ((local / kXcdChunk) * kXcds + x) * kXcdChunk + local % kXcdChunk;
}

// ---- one sub-cell ray ------------------------------------------------------------------

// Per sub-cell quantities shared by every lane (detector-side geometry).
struct CellCtx {
    double q[3];      // sub-cell centre, world
    double qs[3];     // sub-cell centre, frame of bar d
    double lo[3], hi[3];  // sub-cell box, frame of bar d
    double area[3];   // projected cell-face areas: x, y, z faces
    float hc[3];      // sub-cell half sizes
    uint32_t gate_clear = 0;  // bit m: no ray of the tile can be stopped by the gate of layer m
};

SMG_FN void make_cell_ctx(const Bar<double>& bd, int a, int b, int c, int nsx, int nsy, int nsz,
                          CellCtx& cc) {
    sub_cell(bd, a, b, c, nsx, nsy, nsz, cc.qs, cc.lo, cc.hi);
    from_frame(bd, cc.qs, cc.q);
    cc.area[0] = bd.t * bd.h / nsz / nsy;
    cc.area[1] = bd.w * bd.h / nsx / nsz;
    cc.area[2] = bd.w * bd.t / nsx / nsy;
    for (int k = 0; k < 3; ++k) cc.hc[k] = float(0.5 * (cc.hi[k] - cc.lo[k]));
}


// Plate gate of a whole tile (PESysMatGen.cu:1240-1254 blocks a ray whose
// chord through a plate is below 0.1).  True when every segment from a point
// of the ball (tc, rho) to q crosses the plate's two y faces inside its x and
// z extents, strictly between its end points: its chord is then at least the
// plate thickness, here at least 0.11, and the gate cannot trigger.  Ray
// points are q + s (tc + e - q), |e| <= rho; the y faces are met at
// s = (Y - q_y) / (u_y + e_y), bounded over e_y = +-rho, and x, z over that s
// range are bounded at its ends.  Margins of 0.01 cover the reference's
// direction-component branches (a component below 0.001 moves the ray less
// than 0.001 along that axis).
SMG_FN bool plate_gate_clear(const Layer<double>& L, const double q[3], const double tc[3],
                             double rho) {
    const double m = 0.01;
    if (!(L.hi[1] - L.lo[1] >= 0.11)) return false;
    const double u[3] = {tc[0] - q[0], tc[1] - q[1], tc[2] - q[2]};
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (!(tabs(u[1]) > rho + 1e-6) || !(lu > rho + 1.0)) return false;
    double s_lo = 1e300, s_hi = -1e300;
    for (int f = 0; f < 2; ++f)
        for (int e = -1; e <= 1; e += 2) {
            const double s = ((f ? L.hi[1] : L.lo[1]) - q[1]) / (u[1] + e * rho);
            s_lo = tmin2(s_lo, s);
            s_hi = tmax2(s_hi, s);
        }
    // strictly inside the segment: more than 0.01 from either end point
    const double ds = m / (lu - rho);
    if (!(s_lo > ds) || !(s_hi < 1.0 - ds)) return false;
    for (int k = 0; k < 3; k += 2) {
        const double a = tmin2(s_lo * (u[k] - rho), s_hi * (u[k] - rho));
        const double z = tmax2(s_lo * (u[k] + rho), s_hi * (u[k] + rho));
        if (!(q[k] + a > L.lo[k] + m) || !(q[k] + z < L.hi[k] - m)) return false;
    }
    return true;
}

// The same test for every segment from a point of the ball (qc, rb) to a
// point of the ball (tc, rho): true when each such segment crosses the
// plate's two y faces inside its x and z extents, more than 0.01 from either
// end point.  With endpoints qc + f and tc + e (|f| <= rb, |e| <= rho) the
// y faces are met at s = (Y - qc_y - f_y) / (u_y + e_y - f_y), u = tc - qc;
// with numerator and denominator of fixed sign, s is monotone in each, so its
// range is spanned by their extremes.  A point at s lies within
// (1 - s) rb + s rho of qc + s u, a bound linear in s, so the x and z bounds
// over an s range are taken at its ends.  Applied to a whole bar (qc its
// centre, rb half its diagonal), it clears the gate for every sub-cell point
// of the bar at once; rb = 0 gives the conditions of plate_gate_clear.
SMG_FN bool plate_gate_clear_ball(const Layer<double>& L, const double qc[3], double rb,
                                  const double tc[3], double rho) {
    const double m = 0.01;
    if (!(L.hi[1] - L.lo[1] >= 0.11)) return false;
    const double u[3] = {tc[0] - qc[0], tc[1] - qc[1], tc[2] - qc[2]};
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double R = rho + rb;
    if (!(tabs(u[1]) > R + 1e-6) || !(lu > R + 1.0)) return false;
    double s_lo = 1e300, s_hi = -1e300;
    for (int f = 0; f < 2; ++f) {
        const double num = (f ? L.hi[1] : L.lo[1]) - qc[1];
        if (!(tabs(num) > rb + 1e-6)) return false;  // a face within reach of the bar ball
        for (int en = -1; en <= 1; en += 2)
            for (int ed = -1; ed <= 1; ed += 2) {
                const double s = (num + en * rb) / (u[1] + ed * R);
                s_lo = tmin2(s_lo, s);
                s_hi = tmax2(s_hi, s);
            }
    }
    // strictly inside the segment: more than 0.01 from either end point (the
    // segment is at least lu - R long)
    const double ds = m / (lu - R);
    if (!(s_lo > ds) || !(s_hi < 1.0 - ds)) return false;
    for (int k = 0; k < 3; k += 2) {
        // qc_k + s u_k -+ ((1 - s) rb + s rho)
        const double a = tmin2(qc[k] - rb + s_lo * (u[k] + rb - rho), qc[k] - rb + s_hi * (u[k] + rb - rho));
        const double z = tmax2(qc[k] + rb + s_lo * (u[k] - rb + rho), qc[k] + rb + s_hi * (u[k] - rb + rho));
        if (!(a > L.lo[k] + m) || !(z < L.hi[k] - m)) return false;
    }
    return true;
}

// Sub-cell number `cell` in the reference's loop order: z outermost, then x,
// then y (PESysMatGen.cu:1193-1197).
SMG_FN void make_cell_ctx_linear(const Bar<double>& bd, int cell, const int sub[3], CellCtx& cc) {
    const int b = cell % sub[1], a = (cell / sub[1]) % sub[0], c = cell / (sub[1] * sub[0]);
    make_cell_ctx(bd, a, b, c, sub[0], sub[1], sub[2], cc);
}

// Candidate lists of one lane.  Provides the bars and holes the ray must be
// tested against; on(j) tells whether candidate j is in this element's
// reference culling set.
//   int nbars() const; int bar(int j) const; bool bar_on(int j) const;
//   int nholes() const; int hole(int j) const; bool hole_on(int j) const;
//   static constexpr bool kFastFaces;  with it, bool fast(int j) const and
//   const float* faces(int j) const: bar j qualifies for fast_bar_chord and
//   its faces relative to the sub-cell point (fast_faces).

// Chord of the ray voxel p -> sub-cell point q through another bar k
// (PESysMatGen.cu:1290-1325).  FP32: slab parameters relative to the point of
// the ray closest to the bar centre, placed in FP64; FP64: the reference form.
template <class T>
SMG_FN T other_bar_chord(const PeGeom& G, int k, const double p[3], const double D[3], double dd,
                         double inv_dd, const double q[3], const SegDir<T>& sw, T tol_d,
                         bool& near) {
    const Bar<double>& bk = G.bars_d[k];
    if constexpr (std::is_same<T, float>::value) {
        const Bar<float>& bf = G.bars_f[k];
        double rel[3] = {bk.cx - p[0], bk.cy - p[1], bk.cz - p[2]};
        double proj = rel[0] * D[0] + rel[1] * D[1] + rel[2] * D[2];
        double t0 = proj * inv_dd;
        double s0 = t0 * inv_dd;
        double o[3] = {D[0] * s0 - rel[0], D[1] * s0 - rel[1], D[2] * s0 - rel[2]};
        float lo[3] = {-0.5f * bf.w, -0.5f * bf.t, -0.5f * bf.h};
        float hi[3] = {0.5f * bf.w, 0.5f * bf.t, 0.5f * bf.h};
        float lim_hi = float(dd - t0), lim_lo = float(double(kSlabEps<float>) - t0);
        if (bk.sb == 0.0 && bk.cb == 1.0) {
            float of[3] = {float(o[0]), float(o[1]), float(o[2])};
            return seg_box_local(sw, of, lim_lo, lim_hi, lo, hi, near);
        }
        double cb = bk.cb, sb = -bk.sb;
        float of[3] = {float(o[0] * cb - o[2] * sb), float(o[1]), float(o[0] * sb + o[2] * cb)};
        SegDir<float> sr;
        segdir_init<float>(sr, float(D[0] * cb - D[2] * sb), float(D[1]), float(D[0] * sb + D[2] * cb),
                           tol_d, near);
        return seg_box_local(sr, of, lim_lo, lim_hi, lo, hi, near);
    } else {
        double a[3], b[3];
        to_frame(bk, p, a);
        to_frame(bk, q, b);
        double lo[3] = {-0.5 * bk.w, -0.5 * bk.t, -0.5 * bk.h};
        double hi[3] = {0.5 * bk.w, 0.5 * bk.t, 0.5 * bk.h};
        SegDir<double> sr;
        segdir_init<double>(sr, b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0, near);
        return seg_box<double>(sr, a, lo, hi, near);
    }
}

// ---- FP32 chords through unturned bars measured from the sub-cell point ---------------
//
// For an unturned bar k and the sub-cell point q, the box faces relative to q
// (formed once per sub-cell and candidate, in FP64, rounded once) give the
// slab parameters of every lane's ray measured from q: t = -(face - q) len /
// D_k (D = q - p).  Each is a product of two correctly rounded factors, so
// its relative error is a few ulp whatever the ray's angle, and the chord
// t_out - t_in of a bar near q (the detector bars are near q, the voxels far)
// keeps ~1e-7 mm.  A bar qualifies when it is more than 0.01 mm from q and
// from the tile ball: then the line of a lane's ray meets it (if at all) in
// an interval entirely in front of q (t_in >= 0.01), entirely behind q
// (t_out <= -0.01: the reference's entry and exit lie past the segment's end,
// chord 0) or entirely beyond the voxel (t_in >= len + 0.01), so the
// reference's tests of the entry and exit against the segment's ends
// (PESysMatGen.cu:297-304) are decided by the sign of t_in and of
// len - t_in with a margin of 0.01, and the chord below eps is the only fine
// decision (PESysMatGen.cu:291), taken with the tolerance of the two
// parameters.

// Faces of unturned bar k relative to q: out = (lo - q, hi - q) per axis;
// true when the bar qualifies (unturned, more than 0.01 mm from q and from
// the ball (tc, rho)).
// The part of the fast_faces test that depends on the tile only: bar k is
// unturned and its box lies more than rho + 0.01 from the tile ball's centre
// (distances compared squared).
SMG_FN bool fast_tile_far(const PeGeom& G, int k, const double tc[3], double rho) {
    const Bar<double>& b = G.bars_d[k];
    const double c[3] = {b.cx, b.cy, b.cz}, h[3] = {0.5 * b.w, 0.5 * b.t, 0.5 * b.h};
    double dt = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double lo = c[i] - h[i], hi = c[i] + h[i];
        const double et = tmax2(tmax2(lo - tc[i], tc[i] - hi), 0.0);
        dt += et * et;
    }
    return b.sb == 0.0 && b.cb == 1.0 && dt > (rho + 0.01) * (rho + 0.01);
}

// The part that depends on the sub-cell point q: the faces relative to q, and
// true when the box lies more than 0.01 from q.
SMG_FN bool fast_faces_q(const PeGeom& G, int k, const double q[3], float out[6]) {
    const Bar<double>& b = G.bars_d[k];
    const double c[3] = {b.cx, b.cy, b.cz}, h[3] = {0.5 * b.w, 0.5 * b.t, 0.5 * b.h};
    double dq = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double lo = c[i] - h[i], hi = c[i] + h[i];
        out[i] = float(lo - q[i]);
        out[3 + i] = float(hi - q[i]);
        const double eq = tmax2(tmax2(lo - q[i], q[i] - hi), 0.0);
        dq += eq * eq;
    }
    return dq > 1e-4;
}

SMG_FN bool fast_faces(const PeGeom& G, int k, const double q[3], const double tc[3], double rho,
                       float out[6]) {
    const bool near_q_ok = fast_faces_q(G, k, q, out);
    return fast_tile_far(G, k, tc, rho) && near_q_ok;
}

// Chord of the ray through a qualifying bar from its faces relative to q
// (length_box_ray, PESysMatGen.cu:236-312, for a segment with no direction
// component below eps).
SMG_FN float fast_bar_chord(const SegDir<float>& s, const float n[6], bool& near) {
    float t_in = -INFINITY, t_out = INFINITY, e = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = -n[k] * s.inv[k], b = -n[3 + k] * s.inv[k];
        t_in = tmax2(t_in, tmin2(a, b));
        t_out = tmin2(t_out, tmax2(a, b));
    }
    const float chord = t_out - t_in;
    if (!(chord > 0.0f) || !(t_in > 0.0f) || !(t_in < s.len)) return 0.0f;
    const float tol = 4.0f * kUnit<float> * (tabs(t_in) + tabs(t_out)) + 1e-9f;
    if (g_lt(chord, kSlabEps<float>, tol, near)) return 0.0f;
    return chord;
}

// World-frame direction data of the ray voxel -> sub-cell point.
template <class T>
SMG_FN void ray_dir(const double p[3], const double q[3], double D[3], double& dd, SegDir<T>& sw,
                    T& tol_d, bool& near) {
    for (int k = 0; k < 3; ++k) D[k] = q[k] - p[k];
    dd = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    tol_d = T(4) * kUnit<T> * (tabs(T(D[0])) + tabs(T(D[1])) + tabs(T(D[2])));
    segdir_init<T>(sw, T(D[0]), T(D[1]), T(D[2]), tol_d, near);
}

// Chords a caller may keep from pe_cell_ray (the adjoint's second use).
constexpr int kKeepChords = 8;

// kEll: the elliptical-hole chord is compiled in (the all-FP64 kernels, which
// the library runs whenever a hole is elliptical); otherwise every hole is
// traced as the parameter file's circle, which keeps the ellipse's code, and
// its registers, out of the FP32 kernels.
// tan_bar (optional): a direction over the bars' mu_total; *tan_att then
// receives sum_k tan_bar[k] l_k over the other bars on the ray (the tangent of
// their attenuation).
template <class T, bool kEll, class Cands>
SMG_FN T pe_cell_ray(const PeGeom& G, int d, const double p[3], const double ps[3],
                     const CellCtx& cc, const Cands& C, bool& near, double* dl_layer = nullptr,
                     double* aux = nullptr, float* keep = nullptr, int keep_stride = 0,
                     const double* tan_bar = nullptr, double* tan_att = nullptr) {
    const GeomDesc& g = G.g;
    const Bar<double>& bd = G.bars_d[d];
    constexpr bool kF32 = std::is_same<T, float>::value;
    double D[3] = {cc.q[0] - p[0], cc.q[1] - p[1], cc.q[2] - p[2]};
    double dsq = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    // FP64 path: |D| and cos(alpha) as the reference forms them; the FP32
    // path takes |D| and 1/|D| from a refined hardware reciprocal square root
    // below, where the hole and bar chords need them
    double dd = 0.0, inv_dd = 0.0, cosang = 0.0;
    if constexpr (!kF32) {
        dd = sqrt(dsq);
        inv_dd = 1.0 / dd;
        cosang = D[1] / dd;
    }

    // solid angle: largest of the three projected cell faces (PESysMatGen.cu:1223-1231)
    T om;
    if constexpr (std::is_same<T, double>::value) {
        const double cb = bd.cb, sb = bd.sb, four_pi = 4.0 * 3.14159265358979323846;
        double omx = cc.area[0] / (four_pi * dsq) * tabs(D[0] * cb - D[2] * sb) / dd;
        double omz = cc.area[2] / (four_pi * dsq) * tabs(D[0] * sb + D[2] * cb) / dd;
        double omy = cc.area[1] / (four_pi * dsq) * tabs(D[1]) / dd;
        om = tmax2(tmax2(omx, omy), omz);
    } else {
        // one reciprocal of 4 pi |D|^3 for the three faces
        const float cb = float(bd.cb), sb = float(bd.sb);
        const float Dx = float(D[0]), Dy = float(D[1]), Dz = float(D[2]);
        const float dsqf = float(dsq);
        const float r = trcp(float(4.0 * 3.14159265358979323846) * dsqf * tsqrt(dsqf));
        float omx = float(cc.area[0]) * tabs(Dx * cb - Dz * sb);
        float omz = float(cc.area[2]) * tabs(Dx * sb + Dz * cb);
        float omy = float(cc.area[1]) * tabs(Dy);
        om = tmax2(tmax2(omx, omy), omz) * r;
    }

    double att = 0.0;
    // world-frame direction data, shared by the plates and unrotated bars
    SegDir<T> sw;
    T tol_d = T(4) * kUnit<T> * (tabs(T(D[0])) + tabs(T(D[1])) + tabs(T(D[2])));
    segdir_init<T>(sw, T(D[0]), T(D[1]), T(D[2]), tol_d, near);
    const T pw[3] = {T(p[0]), T(p[1]), T(p[2])};

    // plate gate: a ray that does not cross a plate is blocked (PESysMatGen.cu:1240-1254)
    for (int m = 0; m < g.n_layers; ++m) {
        if ((cc.gate_clear >> m) & 1u) continue;
        T lo[3], hi[3];
        if constexpr (std::is_same<T, float>::value) {
            const Layer<float>& L = G.layers_f[m];
            for (int k = 0; k < 3; ++k) {
                lo[k] = L.lo[k];
                hi[k] = L.hi[k];
            }
        } else {
            const Layer<double>& L = G.layers_d[m];
            for (int k = 0; k < 3; ++k) {
                lo[k] = L.lo[k];
                hi[k] = L.hi[k];
            }
        }
        T l = seg_box<T>(sw, pw, lo, hi, near);
        if (g_lt(l, T(0.1f), T(1e-4), near)) att += 100000.0;
    }

    // holes and plate material, FP64 (PESysMatGen.cu:1257-1286).  The FP32
    // path forms T / cos(alpha) as T |D| / D_y and the hole-ray set-up from
    // the same two reciprocals, only when the sub-cell has hole candidates.
    CylRay cr;
    double inv_cos;
    if constexpr (std::is_same<T, double>::value) {
        cyl_ray(D, dd, cr);
        inv_cos = 0.0;
    } else {
        inv_dd = rsq_fast(dsq);
        dd = dsq * inv_dd;
        inv_cos = tabs(D[1]) > 1e-300 ? dd * rcp_fast(D[1]) : dd / D[1];
        if (C.nholes() > 0) cyl_ray_from(D, inv_dd, inv_cos, cr);
    }
    for (int m = 0; m < g.n_layers; ++m) {
        const Layer<double>& L = G.layers_d[m];
        int first = g.ranges[m].first, last = first + g.ranges[m].count;
        double in_coll;
        if constexpr (std::is_same<T, double>::value)
            in_coll = L.thick / cosang;
        else
            in_coll = L.thick * inv_cos;
        double sum = 0.0;
        for (int j = 0; j < C.nholes(); ++j) {
            int h = C.hole(j);
            if (h < first || h >= last || !C.hole_on(j)) continue;
            const Hole<double>& o = G.holes_d[h];
            double l;
            if constexpr (kEll)
                l = o.erx > 0.0 ? ell_chord(o, p[0], p[1], p[2], cr.kx, cr.kz, cr.inv_ky, dd)
                                : cyl_chord(p, cr, dd, o.xc, o.y1, o.y2, o.zc, o.r);
            else
                l = cyl_chord<kF32>(p, cr, dd, o.xc, o.y1, o.y2, o.zc, o.r);
            sum += l;
            att += o.mu * l;
        }
        double rest = in_coll - sum;
        bool counted = rest >= (double)0.00001f;
        if (counted) att += L.mu * rest;
        if (dl_layer) store_at<kMaxLayers>(dl_layer, m, counted ? rest : 0.0);
    }

    // other bars on the way (PESysMatGen.cu:1290-1325); with `keep`, the
    // chords of the first kKeepChords candidates are also stored (0 where the
    // candidate is not on this lane's list), keep[j * keep_stride]
    for (int j = 0; j < C.nbars(); ++j) {
        if (!C.bar_on(j)) {
            if (keep && j < kKeepChords) keep[j * keep_stride] = 0.0f;
            continue;
        }
        T l;
        double mu_k;
        if constexpr (Cands::kFastFaces && std::is_same<T, float>::value) {
            // the first narrowed bars carry their mu_total in the lists: a
            // fast chord then reads no bar record
            if (C.has_mu(j)) {
                mu_k = double(C.mu(j));
                if (sw.ns == 0 && C.fast(j))
                    l = fast_bar_chord(sw, C.faces(j), near);
                else
                    l = other_bar_chord<T>(G, C.bar(j), p, D, dd, inv_dd, cc.q, sw, tol_d, near);
            } else {
                const int k = C.bar(j);
                mu_k = G.bars_d[k].mut;
                l = other_bar_chord<T>(G, k, p, D, dd, inv_dd, cc.q, sw, tol_d, near);
            }
        } else {
            const int k = C.bar(j);
            mu_k = G.bars_d[k].mut;
            l = other_bar_chord<T>(G, k, p, D, dd, inv_dd, cc.q, sw, tol_d, near);
        }
        if (keep && j < kKeepChords) keep[j * keep_stride] = float(l);
        att += double(l) * mu_k;
        if (tan_bar) *tan_att += double(l) * tan_bar[C.bar(j)];
    }

    // own bar from its surface to the sub-cell (PESysMatGen.cu:1329-1346), and the
    // absorption chord along the ray extended 1000 mm (PESysMatGen.cu:1350-1357)
    double Ds[3] = {cc.qs[0] - ps[0], cc.qs[1] - ps[1], cc.qs[2] - ps[2]};
    T l1, l2, la;
    if constexpr (std::is_same<T, float>::value) {
        // The sub-cell centre q lies inside bar d and is the centre of its
        // sub-cell, and the voxel lies outside bar d.  Then the reference's
        // chords (PESysMatGen.cu:1329-1357) reduce to distances from q:
        //   l1 = len - t_entry(bar)  = min_k dist_k(q to the bar face on the voxel's side) / |u_k|
        //   l2 = len - t_entry(cell) = min_k hc_k / |u_k|
        //   la = t_exit - t_entry of the extended ray through the cell = 2 l2
        // with u the unit direction over the axes whose component is not below
        // eps (the reference's one- and two-small-component branches,
        // PESysMatGen.cu:333-434, drop those axes; the premises put the start
        // strictly inside their faces, as the branches require); no
        // difference of two ray parameters is formed.  Rays outside the
        // premises (all components below eps, |D_y| near 0.001, the voxel
        // within 0.01 of bar d along the ray, sub-cells below 0.01) are
        // flagged for the FP64 path, which keeps the reference's general form.
        const Bar<float>& bf = G.bars_f[d];
        SegDir<float> ss;
        float tds = 4.0f * kUnit<float> * (tabs(float(Ds[0])) + tabs(float(Ds[1])) + tabs(float(Ds[2])));
        segdir_init<float>(ss, float(Ds[0]), float(Ds[1]), float(Ds[2]), tds, near);
        near |= ss.ns == 3;
        const float hb[3] = {0.5f * bf.w, 0.5f * bf.t, 0.5f * bf.h};
        float d1 = INFINITY, d2 = INFINITY;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (ss.small[k]) continue;
            const float ia = tabs(ss.inv[k]);  // 1 / |u_k|
            const float qk = float(cc.qs[k]);
            const float face = ss.D[k] > 0.0f ? qk + hb[k] : hb[k] - qk;
            d1 = tmin2(d1, face * ia);
            d2 = tmin2(d2, cc.hc[k] * ia);
        }
        const bool flat = g_lt(tabs(ss.D[1]), float(kInsideDy), ss.tol_d, near);
        near |= (ss.len - d1 < 0.01f) | (d2 < 0.01f);
        l1 = flat ? 0.0f : d1;
        l2 = flat ? 0.0f : d2;
        la = 2.0f * d2;
    } else {
        double ext[3], De[3];
        for (int k = 0; k < 3; ++k) {
            ext[k] = cc.qs[k] + 1000.0 * (cc.qs[k] - ps[k]) / dd;
            De[k] = ext[k] - ps[k];
        }
        double blo[3] = {-bd.w / 2.0, -bd.t / 2.0, -bd.h / 2.0};
        double bhi[3] = {bd.w / 2.0, bd.t / 2.0, bd.h / 2.0};
        SegDir<double> ss, se;
        segdir_init<double>(ss, Ds[0], Ds[1], Ds[2], 0.0, near);
        l1 = seg_box_inside<double>(ss, ps, blo, bhi, near);
        l2 = seg_box_inside<double>(ss, ps, cc.lo, cc.hi, near);
        segdir_init<double>(se, De[0], De[1], De[2], 0.0, near);
        la = seg_box<double>(se, ps, cc.lo, cc.hi, near);
    }
    att += (double(l1) - double(l2)) * bd.mut;
    T eff = exp_neg<T>(att);
    T mut = T(bd.mut), mupe = T(bd.mupe);
    T ea = exp_neg<T>(double(la) * bd.mut);
    T absf = (T(1) - ea) * T(bd.pe_frac);
    if (aux) {  // quantities the adjoint needs
        aux[0] = double(eff) * double(om);  // Omega e^-att
        aux[1] = double(l1) - double(l2);   // own-bar path
        aux[2] = double(la);                // absorption chord
        aux[3] = double(ea);
        aux[4] = double(l2);                // the sub-cell's own chord
        aux[5] = double(eff);               // e^-att
        aux[6] = dsq;                       // |q - p|^2
    }
    return eff * om * absf;
}

}  // namespace smg
