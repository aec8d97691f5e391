// Primary-Compton scatter matrices for gfx950: crystal scatter (scatter point
// = centre of another bar, weighted by the PE matrix) and collimator scatter
// (scatter point = centre of a layer-0 hole, weighted by a virtual-disc solid
// angle).  References: ScatterGen_RayTracing_CircularHole/scatter.cu:1077-1437
// and :1544-1936, ScatterGen_Collimator/scatter.cu:1056-1435.
//
// Two kernels:
//   pair_terms_kernel  one thread per (bar d, scatter point s, sub-cell): the
//                      voxel-independent second-leg sums S_pe, S_c over the
//                      bars flagged in the geometry bitmap, the own-bar path
//                      and the absorption chord, in FP64.
//   scatter_kernel     one wave per (bar d, 64 consecutive voxels[, slice of
//                      the scatter points]); each lane owns one element and
//                      loops over scatter points and sub-cells, summing with
//                      a two-sum compensated sum.  When the points are split
//                      into slices, scatter_reduce_kernel adds the slices'
//                      sums in slice order.
//                      The reference instead runs one thread per
//                      (d, v, s) and sums with float atomicAdd, in an order
//                      that changes from run to run (scatter.cu:1427).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "smg_kernels.hpp"
#include "smg_scatter_body.hpp"

namespace smg {

namespace {

// The energy-window cut of bar d's row (window_reject_cos with no margin).
__device__ double row_window_cos(const ScatterLaunch& A, const Bar<double>& bd) {
    ScatterPhys ph;
    ph.e0 = A.e0;
    if (A.same_window) {
        ph.lo = A.win_lo;
        ph.hi = A.win_hi;
    } else {  // per-bar window E0 (1 -+ res/2) (scatter.cu:1154-1163)
        ph.lo = (1 - bd.res / 2.0) * A.e0;
        ph.hi = (1 + bd.res / 2.0) * A.e0;
    }
    ph.res = bd.res;
    return window_reject_cos(ph, 0.0);
}

// One thread per output row: its energy-window cut, once per launch instead
// of once per pair-term thread and per scatter wave (an 80-step FP64
// bisection).
__global__ __launch_bounds__(64) void row_cut_kernel(ScatterLaunch A, double* cut) {
    const int r = (int)(blockIdx.x * 64 + threadIdx.x);
    const int n_rows = A.rows ? A.n_rows : A.geom.g.n_det;
    if (r >= n_rows) return;
    cut[r] = row_window_cos(A, A.geom.bars_d[A.rows ? A.rows[r] : r]);
}

// True when the terms of scatter point qs and sub-cell centre q lie below the
// energy window for every voxel of the image, so scatter_kernel's per-lane
// test (below_window, cosine threshold c_lo - 1e-5 in FP32, c_lo - 1e-9 in
// FP64) rejects each of them before it reads the pair term; c_lo is the row's
// cut (row_window_cos).  The voxel centres lie in a ball (C, R); seen from qs
// they are within asin(R / |qs - C|) of the direction qs - C, so no voxel's
// cosine exceeds cos(angle(qs - C, q - qs) - asin(R / |qs - C|)); the test
// asks that bound to be 1e-4 below the FP32 threshold.
__device__ bool below_window_all(const ScatterLaunch& A, double c_lo, const double qs[3],
                                 const double q[3]) {
    if (c_lo <= -1.5) return false;  // nothing is rejected
    const GeomDesc& g = A.geom.g;
    const int whole[3] = {g.nx, g.ny, g.nz};
    double C[3], R;
    tile_ball(g, 0, 0, 0, whole, A.cphi, A.sphi, C, R);
    double u[3], b[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = qs[k] - C[k];
        b[k] = q[k] - qs[k];
    }
    const double L = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double B = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (!(L > R * (1.0 + 1e-9) + 1e-9) || !(B > 0.0)) return false;
    double cu = (u[0] * b[0] + u[1] * b[1] + u[2] * b[2]) / (L * B);
    cu = cu > 1.0 ? 1.0 : (cu < -1.0 ? -1.0 : cu);
    const double th = acos(cu) - asin(R / L);
    if (!(th > 0.0)) return false;
    return cos(th) < c_lo - 1.1e-4;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void pair_terms_kernel(ScatterLaunch A) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const PeGeom& G = A.geom;
    const int N = G.g.n_det;
    const int ncell = A.sub[0] * A.sub[1] * A.sub[2];
    const int n_rows = A.rows ? A.n_rows : N;
    const int64_t total = (int64_t)n_rows * A.n_pts * ncell;
    if (tid >= total) return;
    const int cell = (int)(tid % ncell);
    const int s = (int)((tid / ncell) % A.n_pts);
    const int r = (int)(tid / ((int64_t)ncell * A.n_pts));
    const int d = A.rows ? A.rows[r] : r;
    // the bar holding the scatter point (crystal scatter)
    const int sbar = A.kind == 3 ? A.pt_bar0 + s / A.pts_per_bar : s;
    PairTerm out{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if ((A.kind == 0 || A.kind == 3) && sbar == d) {
        A.pair[tid] = out;
        return;
    }
    const Bar<double>& bd = G.bars_d[d];
    const int a = cell % A.sub[0], b = (cell / A.sub[0]) % A.sub[1], c = cell / (A.sub[0] * A.sub[1]);
    CellCtx cc;
    make_cell_ctx(bd, a, b, c, A.sub[0], A.sub[1], A.sub[2], cc);
    const ScatterPt<double>& sp = A.pts_d[s];
    double qs[3] = {sp.x, sp.y, sp.z};
    if (below_window_all(A, A.row_cut[r], qs, cc.q)) {  // never read: every voxel's term is below
        A.pair[tid] = out;                    // the window (scatter_kernel tests first)
        return;
    }
    bool near = false;
    // bars flagged between the scatter point and bar d (scatter.cu:1343-1401);
    // grid scatterers: every other bar
    int64_t base = 0;
    if (A.kind == 0 || A.kind == 3) {
        int64_t i = sbar < d ? sbar : d, j = sbar < d ? d : sbar;
        base = (i * N - i * (i + 1) / 2 + j) * N;
    } else if (A.kind == 1) {
        base = ((int64_t)s * N + d) * N;
    }
    // In the frame of an unturned bar (beta = 0) the segment's direction is
    // q - qs, the same for every such bar: its direction data (length, the
    // eps decisions, len / D_k) are formed once per record.  A turned bar
    // gets its own.
    SegDir<double> sdw;
    segdir_init<double>(sdw, cc.q[0] - qs[0], cc.q[1] - qs[1], cc.q[2] - qs[2], 0.0, near);
    auto add_bar = [&](int k) {
        const Bar<double>& bk = G.bars_d[k];
        double lo[3] = {-0.5 * bk.w, -0.5 * bk.t, -0.5 * bk.h};
        double hi[3] = {0.5 * bk.w, 0.5 * bk.t, 0.5 * bk.h};
        double l;
        if (bk.sb == 0.0 && bk.cb == 1.0) {
            const double pa[3] = {qs[0] - bk.cx, qs[1] - bk.cy, qs[2] - bk.cz};
            l = seg_box<double>(sdw, pa, lo, hi, near);
        } else {
            double pa[3], pb[3];
            to_frame(bk, qs, pa);
            to_frame(bk, cc.q, pb);
            SegDir<double> sd;
            segdir_init<double>(sd, pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2], 0.0, near);
            l = seg_box<double>(sd, pa, lo, hi, near);
        }
        out.s_pe += l * bk.mupe;
        out.s_c += l * bk.muc;
    };
    if (A.kind == 2) {
        for (int k = 0; k < N; ++k)
            if (k != d) add_bar(k);
    } else {
        // the set bits of bits base .. base + N - 1, a word at a time, in
        // increasing k (the order of a bit-by-bit scan)
        const int64_t w0 = base >> 5, w1 = (base + N - 1) >> 5;
        for (int64_t wi = w0; wi <= w1; ++wi) {
            uint32_t bits = A.bitmap[wi];
            if (wi == w0) bits &= ~0u << (base & 31);
            if (wi == w1) bits &= ~0u >> (31 - ((base + N - 1) & 31));
            while (bits) {
                const int k = (int)(wi * 32 + __builtin_ctz(bits) - base);
                bits &= bits - 1;
                if (k == d || ((A.kind == 0 || A.kind == 3) && k == sbar)) continue;
                add_bar(k);
            }
        }
    }
    if (A.kind == 2) {  // collimator material from the grid point to the sub-cell
        double att_t;
        coll_material(G, qs, cc.q, att_t, out.s_wpe, out.s_wc);
    }
    if (A.kind == 3) {  // grid point inside bar sbar: its way out of that bar
        const Bar<double>& bs = G.bars_d[sbar];
        double sa[3], qa[3];
        to_frame(bs, qs, sa);
        to_frame(bs, cc.q, qa);
        const double lo[3] = {-0.5 * bs.w, -0.5 * bs.t, -0.5 * bs.h};
        const double hi[3] = {0.5 * bs.w, 0.5 * bs.t, 0.5 * bs.h};
        const double l = seg_box_geo(sa, qa, lo, hi);
        out.s_pe += l * bs.mupe;
        out.s_c += l * bs.muc;
    }
    // own bar from its surface to the sub-cell (scatter.cu:1405-1417)
    double qss[3];
    to_frame(bd, qs, qss);
    double blo[3] = {-bd.w / 2.0, -bd.t / 2.0, -bd.h / 2.0};
    double bhi[3] = {bd.w / 2.0, bd.t / 2.0, bd.h / 2.0};
    SegDir<double> so;
    segdir_init<double>(so, cc.qs[0] - qss[0], cc.qs[1] - qss[1], cc.qs[2] - qss[2], 0.0, near);
    double l1 = seg_box_inside<double>(so, qss, blo, bhi, near);
    double l2 = seg_box_inside<double>(so, qss, cc.lo, cc.hi, near);
    out.s_pe += (l1 - l2) * bd.mupe;
    out.s_c += (l1 - l2) * bd.muc;
    // absorption chord, ray extended 1000 mm through the sub-cell (scatter.cu:1420-1424)
    double dx = qs[0] - cc.q[0], dy = qs[1] - cc.q[1], dz = qs[2] - cc.q[2];
    out.len = sqrt(dx * dx + dy * dy + dz * dz);
    double De[3];
    for (int k = 0; k < 3; ++k) De[k] = (cc.qs[k] + 1000.0 * (cc.qs[k] - qss[k]) / out.len) - qss[k];
    SegDir<double> se;
    segdir_init<double>(se, De[0], De[1], De[2], 0.0, near);
    out.la = seg_box<double>(se, qss, cc.lo, cc.hi, near);
    A.pair[tid] = out;
}

// Waves per element row: the scatter points split into `slices` ranges, each
// range summed by its own wave, when rows x voxel waves alone would leave the
// chip under-filled (a few detector rows, or a small image: up to 16,384
// waves), and into up to 16 ranges towards 262,144 waves otherwise.  The
// waves' costs differ widely (the terms that pass the energy window depend on
// the row and the voxels), so short waves shorten the launch's tail: the
// 0.5 mm collimator grid (16,384 unsliced waves per chunk) ran 12% faster
// with 16 slices, its resident waves per SIMD having averaged 1.9 of 3.
struct SliceShape {
    int64_t rows, waves_per_det;
    int slices, per;  // per: scatter points per slice
};

SliceShape slice_shape(const ScatterLaunch& A) {
    SliceShape s;
    s.rows = A.rows ? A.n_rows : A.geom.g.n_det;
    s.waves_per_det = (A.geom.g.n_vox + 63) / 64;
    const int64_t base = s.rows * s.waves_per_det;
    constexpr int64_t kFillWaves = 16384;     // 256 CUs x 4 SIMDs x 16
    constexpr int64_t kTailWaves = 262144;    // 256 CUs x 4 SIMDs x 256
    constexpr int64_t kTailSlices = 16;
    constexpr int64_t kMinPoints = 32;        // scatter points per wave, at least
    int64_t n = 1;
    if (base > 0)
        n = std::max<int64_t>((kFillWaves + base - 1) / base,
                              std::min<int64_t>(kTailSlices, (kTailWaves + base - 1) / base));
    n = std::min<int64_t>(n, std::max<int64_t>(1, (int64_t)A.n_pts / kMinPoints));
    n = std::max<int64_t>(n, 1);
    s.per = (int)((A.n_pts + n - 1) / n);
    s.slices = s.per > 0 ? (int)((A.n_pts + s.per - 1) / s.per) : 1;
    if (s.slices < 1) s.slices = 1;
    return s;
}

// Per-wave term lists of scatter_kernel, in LDS.  Most terms fail the
// energy-window test, and those that pass are scattered over the lanes, so a
// wave that evaluated the term of every lane for each scatter point in turn
// kept about half its lanes idle.  Instead a batch of scatter points is first
// tested (each lane for its voxel), the (point, lane) pairs that pass are
// listed, the terms of the list are evaluated 64 at a time (lane k takes pair
// k, whatever its voxel), and each lane then adds its own terms in point
// order: the same sums, in the same order, as the per-point loop.
constexpr int kScatterBatchPts = 128;  // points with a passing lane per batch
template <bool kExact>
struct ScatterLists {
    static constexpr int kCap = kExact ? 512 : 1024;  // pairs per batch
    using T = typename std::conditional<kExact, double, float>::type;
    uint64_t mask[kScatterBatchPts];  // lanes whose term passes, per batch point
    int pre[kScatterBatchPts];        // pairs listed before the point
    int pt[kScatterBatchPts];         // the point's index
    uint16_t pair[kCap];              // (batch point << 6) | lane
    T val[kCap];                      // the pairs' terms
    double vp[3][64], vps[3][64];     // the lanes' voxels: world, frame of bar d
};

// 3 waves per SIMD: the register cap (168) spills a few values that the
// inline FP64 fallback holds; measured 14% faster (FP32) and 5% (FP64) than the
// 2 waves the unconstrained allocation gives.  LDS: 11.3 KB (FP32) / 10.2 KB
// (FP64) per wave, 12 waves per CU.
template <bool kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void scatter_kernel(
    ScatterLaunch A, SliceShape sh) {
    using T = typename std::conditional<kExact, double, float>::type;
    using Lists = ScatterLists<kExact>;
    __shared__ Lists S;
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int64_t waves_per_det = sh.waves_per_det;
    const int64_t wrow = blockIdx.x % (sh.rows * waves_per_det);
    const int slice = (int)(blockIdx.x / (sh.rows * waves_per_det));
    const int r = (int)(wrow / waves_per_det);  // output row
    const int d = A.rows ? A.rows[r] : r;
    const int64_t v = (wrow % waves_per_det) * 64 + lane;
    const bool valid = v < g.n_vox;
    const bool sliced = sh.slices > 1;
    float* out = A.dst + (int64_t)r * g.n_vox;
    double* part = A.partial + ((int64_t)slice * sh.rows + r) * g.n_vox;
    const int s_begin = slice * sh.per;
    const int s_end = min(A.n_pts, s_begin + sh.per);
    const Bar<double>& bd = G.bars_d[d];
    if (!A.compton_on || !(bd.mut > (double)0.01f)) {  // scatter.cu:1226-1228
        if (valid) {
            if (sliced)
                part[v] = 0.0;
            else
                out[v] = A.accumulate ? out[v] : 0.0f;
        }
        return;
    }
    const int64_t vv = valid ? v : 0;
    const int nxy = g.nx * g.ny;
    const int iz = (int)(vv / nxy), rem = (int)(vv % nxy);
    double p[3], ps[3];
    voxel_position(g, rem % g.nx, rem / g.nx, iz, A.cphi, A.sphi, p);
    to_frame(bd, p, ps);

    ScatterPhys ph;
    ph.e0 = A.e0;
    if (A.same_window) {
        ph.lo = A.win_lo;
        ph.hi = A.win_hi;
    } else {  // per-bar window E0 (1 -+ res/2) (scatter.cu:1154-1163)
        ph.lo = (1 - bd.res / 2.0) * A.e0;
        ph.hi = (1 + bd.res / 2.0) * A.e0;
    }
    ph.norm = A.kn_norm;
    ph.tab = A.tables;
    ph.tabf = A.tables_f;
    ph.knt = A.kn_table;
    ph.res = bd.res;
    ph.mupe = bd.mupe;
    ph.muc = bd.muc;
    {
        const double sx = A.sub[0], sy = A.sub[1], sz = A.sub[2];
        ph.rcell = sqrt(bd.w * bd.w / sx / sx + bd.h * bd.h / sz / sz + bd.t * bd.t / sy / sy) / 2.0;
    }
    ph.c_rej = window_cut_margin(A.row_cut[r], kExact ? 1e-9 : 1e-5);  // window_reject_cos(ph, .)
    // the row's physics is the same for every lane: scalar registers
    ph.e0 = uniform_d(ph.e0);
    ph.lo = uniform_d(ph.lo);
    ph.hi = uniform_d(ph.hi);
    ph.norm = uniform_d(ph.norm);
    ph.res = uniform_d(ph.res);
    ph.mupe = uniform_d(ph.mupe);
    ph.muc = uniform_d(ph.muc);
    ph.rcell = uniform_d(ph.rcell);
    ph.c_rej = uniform_d(ph.c_rej);
    const int ncell = A.sub[0] * A.sub[1] * A.sub[2];
    const T wscale = T(bd.muc / bd.mupe);  // crystal weight uses bar d's mu (scatter.cu:1218-1219)
    // scatter points inside bar d itself do not count: its centre (kind 0),
    // or its grid points (kind 3)
    int skip_lo = 0, skip_hi = 0;
    if (A.kind == 0) {
        skip_lo = d;
        skip_hi = d + 1;
    } else if (A.kind == 3) {
        skip_lo = (d - A.pt_bar0) * A.pts_per_bar;
        skip_hi = skip_lo + A.pts_per_bar;
    }
    const int64_t v0 = v - lane;  // first voxel of the wave
    for (int k = 0; k < 3; ++k) {
        S.vp[k][lane] = p[k];
        S.vps[k][lane] = ps[k];
    }
    // weight of scatter point s for voxel vo at po
    auto weight = [&](int s, int64_t vo, const double po[3]) -> T {
        if (A.kind == 0) return T(A.pe[(int64_t)s * g.n_vox + vo]) * wscale;
        if (A.kind == 2 || A.kind == 3) return T(A.w1[(int64_t)s * g.n_vox + vo]);
        const ScatterPt<double>& sp = A.pts_d[s];
        double dx = po[0] - sp.x, dy = po[1] - sp.y, dz = po[2] - sp.z;
        double lv = sqrt(dx * dx + dy * dy + dz * dz);
        double om;
        if (A.variant == 0) {  // scatter.cu:1711-1716
            double rr = (double)0.1f;
            om = (rr * rr * 3.14159265358979323846 + sp.dy * rr) /
                 (4.0 * 3.14159265358979323846 * lv * lv) * fabs(sp.y - po[1]) / lv;
        } else {  // ScatterGen_Collimator/scatter.cu:1212-1215
            double rr = (double)0.2f;
            om = (rr * rr + sp.dy * rr) / (4.0 * lv * lv);
        }
        return T(om * A.coll_ratio);
    };
    const uint64_t below = (1ull << lane) - 1ull;  // the lanes under this one
    KahanSum<T> acc;
    unsigned fallbacks = 0;
    for (int cell = 0; cell < ncell; ++cell) {
        const int a = cell % A.sub[0], b = (cell / A.sub[0]) % A.sub[1], c = cell / (A.sub[0] * A.sub[1]);
        CellCtx cc;
        make_cell_ctx(bd, a, b, c, A.sub[0], A.sub[1], A.sub[2], cc);
        ScatterCell sc;
        for (int k = 0; k < 3; ++k) {  // the same for every lane
            sc.q[k] = uniform_d(cc.q[k]);
            sc.lo[k] = uniform_d(cc.lo[k]);
            sc.hi[k] = uniform_d(cc.hi[k]);
        }
        int s = s_begin;
        while (s < s_end) {  // wave-uniform
            // 1. the batch: points in order until kScatterBatchPts of them have
            // a passing lane or the list could overflow with the next point
            int n = 0, nb = 0;
            while (s < s_end && nb < kScatterBatchPts && n <= Lists::kCap - 64) {
                const int si = s++;
                if (si >= skip_lo && si < skip_hi) continue;  // wave-uniform
                const ScatterPt<double>& sp = A.pts_d[si];
                sc.qs[0] = sp.x;
                sc.qs[1] = sp.y;
                sc.qs[2] = sp.z;
                const bool pass = valid && !below_window<T>(ph, p, sc);
                const uint64_t m = __ballot(pass);
                if (m) {  // wave-uniform
                    if (pass) S.pair[n + __popcll(m & below)] = (uint16_t)((nb << 6) | lane);
                    if (lane == 0) {
                        S.mask[nb] = m;
                        S.pre[nb] = n;
                        S.pt[nb] = si;
                    }
                    n += __popcll(m);
                    ++nb;
                }
            }
            __syncthreads();
            // 2. the listed terms, 64 at a time
            for (int base = 0; base < n; base += 64) {
                const int k = base + lane;
                if (k < n) {
                    const int e = S.pair[k];
                    const int j = e >> 6, o = e & 63;
                    const int si = S.pt[j];
                    const double po[3] = {S.vp[0][o], S.vp[1][o], S.vp[2][o]};
                    const double pso[3] = {S.vps[0][o], S.vps[1][o], S.vps[2][o]};
                    ScatterCell sk;
                    for (int q = 0; q < 3; ++q) {
                        sk.q[q] = sc.q[q];
                        sk.lo[q] = sc.lo[q];
                        sk.hi[q] = sc.hi[q];
                    }
                    const ScatterPt<double>& sp = A.pts_d[si];
                    sk.qs[0] = sp.x;
                    sk.qs[1] = sp.y;
                    sk.qs[2] = sp.z;
                    to_frame(bd, sk.qs, sk.qss);
                    sk.pt = A.pair[((int64_t)r * A.n_pts + si) * ncell + cell];
                    const T wgt = weight(si, v0 + o, po);
                    T val = T(0);
                    if (wgt != T(0)) {
                        bool near = false;
                        val = scatter_term<T>(ph, po, pso, sk, wgt, near);
                        if constexpr (!kExact) {
                            if (near) {
                                bool n2 = false;
                                val = T(scatter_term<double>(ph, po, pso, sk, double(wgt), n2));
                                ++fallbacks;
                            }
                        }
                    }
                    S.val[k] = val;
                }
            }
            __syncthreads();
            // 3. each lane's terms in point order (a term of weight 0 adds
            // an exact 0)
            for (int j = 0; j < nb; ++j) {
                const uint64_t m = S.mask[j];
                if ((m >> lane) & 1ull) acc.add(S.val[S.pre[j] + __popcll(m & below)]);
            }
            __syncthreads();  // the lists are refilled by the next batch
        }
    }
    if (valid) {
        if (sliced)
            part[v] = double(acc.s) + double(acc.c);
        else
            out[v] = A.accumulate ? out[v] + float(acc.value()) : float(acc.value());
    }
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
}

// One thread per row of the KN prefix table.
__global__ __launch_bounds__(64) void kn_table_kernel(double alpha, double* tab) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kKnPhase) return;
    kn_prefix_row(alpha, j, tab + (size_t)j * kKnSteps);
}

// Sum of the slices' partial sums of one element, in slice order (a fixed
// order, so the result does not depend on scheduling), compensated in FP64.
__global__ __launch_bounds__(256) void scatter_reduce_kernel(ScatterLaunch A, SliceShape sh) {
    const int64_t n_vox = A.geom.g.n_vox;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = sh.rows * n_vox;
    if (tid >= n) return;
    KahanSum<double> acc;
    for (int k = 0; k < sh.slices; ++k) acc.add(A.partial[(int64_t)k * n + tid]);
    const float val = float(acc.value());
    A.dst[tid] = A.accumulate ? A.dst[tid] + val : val;
}

// Grid scatterers (an extension with no reference counterpart): one thread
// per (scatter point s, voxel v) of the chunk, the chance of a Compton
// interaction in the cell around s for a photon from v,
// w_s / (4 pi |s - v|^2) exp(-collimator material along v -> s), in FP64,
// with w_s = dv mu_c of the point's layer (held in the point's dy).
__global__ __launch_bounds__(256) void first_leg_kernel(ScatterLaunch A, float* w1) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const GeomDesc& g = A.geom.g;
    const int64_t total = (int64_t)A.n_pts * g.n_vox;
    if (tid >= total) return;
    const int s = (int)(tid / g.n_vox);
    const int64_t v = tid % g.n_vox;
    const int nxy = g.nx * g.ny;
    const int iz = (int)(v / nxy), rem = (int)(v % nxy);
    double p[3];
    voxel_position(g, rem % g.nx, rem / g.nx, iz, A.cphi, A.sphi, p);
    const ScatterPt<double>& sp = A.pts_d[s];
    const double q[3] = {sp.x, sp.y, sp.z};
    double att_t, s_pe, s_c;
    coll_material(A.geom, p, q, att_t, s_pe, s_c);
    const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    const double L2 = dx * dx + dy * dy + dz * dz;
    w1[tid] = float(sp.dy / (4.0 * 3.14159265358979323846 * L2) * exp(-att_t));
}

}  // namespace

hipError_t launch_first_leg(const ScatterLaunch& A, float* w1, hipStream_t stream) {
    const int64_t total = (int64_t)A.n_pts * A.geom.g.n_vox;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(first_leg_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       A, w1);
    return hipGetLastError();
}

hipError_t launch_kn_table(double e0, double* tab, hipStream_t stream) {
    hipLaunchKernelGGL(kn_table_kernel, dim3((kKnPhase + 1 + 63) / 64), dim3(64), 0, stream, e0 / 511.0, tab);
    return hipGetLastError();
}

hipError_t launch_row_cuts(const ScatterLaunch& A, double* cut, hipStream_t stream) {
    const int64_t rows = A.rows ? A.n_rows : A.geom.g.n_det;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(row_cut_kernel, dim3((unsigned)((rows + 63) / 64)), dim3(64), 0, stream, A, cut);
    return hipGetLastError();
}

hipError_t launch_pair_terms(const ScatterLaunch& A, hipStream_t stream) {
    const int64_t rows = A.rows ? A.n_rows : A.geom.g.n_det;
    const int64_t total = rows * A.n_pts * A.sub[0] * A.sub[1] * A.sub[2];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(pair_terms_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, A);
    return hipGetLastError();
}

size_t scatter_partial_size(const ScatterLaunch& A) {
    const SliceShape sh = slice_shape(A);
    return sh.slices > 1 ? (size_t)sh.slices * (size_t)sh.rows * (size_t)A.geom.g.n_vox : 0;
}

hipError_t launch_scatter(const ScatterLaunch& A, bool exact, hipStream_t stream) {
    const SliceShape sh = slice_shape(A);
    const int64_t blocks = sh.slices * sh.rows * sh.waves_per_det;
    if (blocks <= 0) return hipSuccess;
    if (sh.slices > 1 && !A.partial) return hipErrorInvalidValue;
    if (exact)
        hipLaunchKernelGGL(scatter_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A, sh);
    else
        hipLaunchKernelGGL(scatter_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A, sh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || sh.slices == 1) return e;
    const int64_t n = sh.rows * A.geom.g.n_vox;
    hipLaunchKernelGGL(scatter_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       A, sh);
    return hipGetLastError();
}

namespace {

// out = a + b, four floats per thread where the three pointers allow it.
__global__ __launch_bounds__(256) void add_blocks_kernel(const float* a, const float* b, float* out,
                                                         size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = a[i] + b[i];
}

__global__ __launch_bounds__(256) void add_blocks4_kernel(const float4* a, const float4* b, float4* out,
                                                          size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 x = a[i], y = b[i];
        out[i] = make_float4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w);
    }
}

}  // namespace

hipError_t launch_add_blocks(const float* a, const float* b, float* out, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const bool al = ((uintptr_t)a | (uintptr_t)b | (uintptr_t)out) % 16 == 0 && n % 4 == 0;
    const size_t work = al ? n / 4 : n;
    const unsigned blocks = (unsigned)std::min<size_t>((work + 255) / 256, 8192);
    if (al)
        hipLaunchKernelGGL(add_blocks4_kernel, dim3(blocks), dim3(256), 0, stream, (const float4*)a,
                           (const float4*)b, (float4*)out, work);
    else
        hipLaunchKernelGGL(add_blocks_kernel, dim3(blocks), dim3(256), 0, stream, a, b, out, n);
    return hipGetLastError();
}

}  // namespace smg
