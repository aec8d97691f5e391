// Photoelectric system-matrix kernel for gfx950 (CDNA4, wave64).
//
// Work decomposition: one wave per (detector bar d, tile of 64 voxels).  The
// reference runs one thread per (d, v) and makes every thread scan all bars
// and holes for culling (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1097-1186)
// before its sub-cell loop (:1193-1362).  Here the 64 lanes of a wave share
// detector d, so every sub-cell point is wave-uniform and:
//   1. the wave scans the bars and holes once, 64 at a time, keeping those
//      whose box a segment from bar d to the tile's bounding ball can meet
//      (any other has chord 0 for every ray of the wave, so whether the
//      reference would cull it does not change the result); each survivor is
//      then tested with the reference culling predicate per lane and stored
//      in LDS with the 64-bit ballot of the lanes that keep it;
//   2. for each sub-cell, the candidates are narrowed to those whose box can
//      meet a segment from the sub-cell centre to the tile's bounding ball (a
//      bar outside that cone has chord 0 for every lane, so dropping it is
//      exact);
//   3. every lane traces its own ray against the narrowed lists (wave-uniform
//      loops, per-lane masks) and accumulates the sub-cell terms with a
//      two-sum compensated sum;
//   4. each lane writes its element once: no atomics, deterministic results.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "smg_kernels.hpp"
#include "smg_pe_body.hpp"

namespace smg {

namespace {

// Capacity of the per-wave queue of sub-cells to redo in FP64.
constexpr int kRedoCap = 64;

// Narrowed bars of a sub-cell whose faces relative to the sub-cell point are
// kept for fast_bar_chord (the first kFastBars).
constexpr int kFastBars = 32;

// Candidate lists of one wave, in LDS, for up to CapB bars and CapH holes.
template <int CapB, int CapH>
struct WaveListsT {
    static constexpr int cap_bars = CapB, cap_holes = CapH;
    uint64_t bar_mask[CapB];
    uint64_t hole_mask[CapH];
    uint64_t tile_far[(CapB + 63) / 64];  // bit i: candidate bar i passes fast_tile_far
    uint16_t bar_idx[CapB];
    uint16_t hole_idx[CapH];
    uint16_t bar_sel[CapB];
    uint16_t hole_sel[CapH];
    float faces[kFastBars][7];   // narrowed bar j: faces - sub-cell point (fast_faces),
                                 // then the bar's mu_total (a float value: exact)
    uint8_t fast_ok[kFastBars];  // narrowed bar j qualifies for fast_bar_chord
};
using WaveLists = WaveListsT<kCapBars, kCapHoles>;

// The adjoint's lists hold fewer bars: with its per-candidate FP64 sums and
// kept chords, 13 KB of LDS per wave, so 12 waves (3 per SIMD) fit a CU.
constexpr int kCapBarsAdjoint = 320;
using AdjointLists = WaveListsT<kCapBarsAdjoint, kCapHoles>;

// Candidate lists of a wave whose tile overflowed the LDS lists, in its own
// slot of global memory with room for every bar and hole of the geometry (so
// they cannot overflow), plus the adjoint's per-candidate sums.
struct SpillLists {
    int cap_bars, cap_holes;
    uint64_t* bar_mask;
    uint64_t* hole_mask;
    uint64_t* tile_far;
    double* cand_acc;
    float (*faces)[7];
    uint16_t* bar_idx;
    uint16_t* hole_idx;
    uint16_t* bar_sel;
    uint16_t* hole_sel;
    uint8_t* fast_ok;
};

__device__ __forceinline__ SpillLists spill_lists(const PeSpill& S, int slot) {
    unsigned char* p = S.slots + (int64_t)slot * S.slot_bytes;
    SpillLists L;
    L.cap_bars = S.cap_bars;
    L.cap_holes = S.cap_holes;
    L.bar_mask = (uint64_t*)p;
    p += sizeof(uint64_t) * S.cap_bars;
    L.hole_mask = (uint64_t*)p;
    p += sizeof(uint64_t) * S.cap_holes;
    L.tile_far = (uint64_t*)p;
    p += sizeof(uint64_t) * ((S.cap_bars + 63) / 64);
    L.cand_acc = (double*)p;
    p += sizeof(double) * S.cap_bars;
    L.faces = (float(*)[7])p;
    p += sizeof(float) * 7 * kFastBars;
    L.bar_idx = (uint16_t*)p;
    p += sizeof(uint16_t) * S.cap_bars;
    L.hole_idx = (uint16_t*)p;
    p += sizeof(uint16_t) * S.cap_holes;
    L.bar_sel = (uint16_t*)p;
    p += sizeof(uint16_t) * S.cap_bars;
    L.hole_sel = (uint16_t*)p;
    p += sizeof(uint16_t) * S.cap_holes;
    L.fast_ok = (uint8_t*)p;
    return L;
}

// A wave whose candidates overflowed its LDS lists queues its work item for
// the spill kernel of the same launch; without a queue (a launch with no
// spill area) it is only counted.
__device__ __forceinline__ void spill_push(const PeLaunch& A, int item, int lane) {
    if (lane == 0) {
        const unsigned long long i = atomicAdd(&A.counters[2], 1ull);
        if (A.spill.queue) A.spill.queue[i] = item;
    }
}

// View of the narrowed lists for one lane.
template <class Lists>
struct ListCands {
    static constexpr bool kFastFaces = true;
    const Lists* L;
    int nb, nh, lane;
    __device__ bool fast(int j) const { return j < kFastBars && L->fast_ok[j]; }
    __device__ const float* faces(int j) const { return L->faces[j]; }
    // mu_total of narrowed bar j, kept beside its faces for the first kFastBars
    __device__ bool has_mu(int j) const { return j < kFastBars; }
    __device__ float mu(int j) const { return L->faces[j][6]; }
    __device__ int nbars() const { return nb; }
    // the index is wave-uniform: readfirstlane lets the record loads go
    // through the scalar unit
    __device__ int bar(int j) const { return __builtin_amdgcn_readfirstlane(L->bar_idx[L->bar_sel[j]]); }
    __device__ bool bar_on(int j) const { return (L->bar_mask[L->bar_sel[j]] >> lane) & 1ull; }
    __device__ int nholes() const { return nh; }
    __device__ int hole(int j) const { return __builtin_amdgcn_readfirstlane(L->hole_idx[L->hole_sel[j]]); }
    __device__ bool hole_on(int j) const { return (L->hole_mask[L->hole_sel[j]] >> lane) & 1ull; }
};

__device__ __forceinline__ int lanes_below(uint64_t m, int lane) {
    return __popcll(m & ((1ull << lane) - 1ull));
}

// Lane and tile state of one wave: voxel of the lane, tile bounding sphere
// and the self-culling terms of bar d.
struct WaveSetup {
    int ix, iy, iz;
    bool valid;
    int64_t v;
    double p[3], ps[3];
    float pf[3];
    SelfCull sc;
    float udf[3], critf;
    double tc[3], rho;
    float tcf[3], rhof;
    uint32_t bar_gate;  // bit m: no ray from bar d to the tile can be stopped by layer m's gate
};

__device__ __forceinline__ void wave_setup(const PeLaunch& A, int d, int tile, int lane, WaveSetup& W) {
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const Bar<double>& bd = G.bars_d[d];
    const int t0 = tile % A.tiles[0];
    const int t1 = (tile / A.tiles[0]) % A.tiles[1];
    const int t2 = tile / (A.tiles[0] * A.tiles[1]);
    const int ix0 = t0 * A.tile[0], iy0 = t1 * A.tile[1], iz0 = t2 * A.tile[2];
    W.ix = ix0 + lane % A.tile[0];
    W.iy = iy0 + (lane / A.tile[0]) % A.tile[1];
    W.iz = iz0 + lane / (A.tile[0] * A.tile[1]);
    W.valid = W.ix < g.nx && W.iy < g.ny && W.iz < g.nz;
    W.v = (int64_t)W.ix + (int64_t)W.iy * g.nx + (int64_t)W.iz * g.nx * g.ny;
    voxel_position(g, W.valid ? W.ix : ix0, W.valid ? W.iy : iy0, W.valid ? W.iz : iz0, A.cphi,
                   A.sphi, W.p);
    to_frame(bd, W.p, W.ps);
    for (int k = 0; k < 3; ++k) W.pf[k] = float(W.p[k]);
    self_cull(bd, W.p, W.sc);
    for (int k = 0; k < 3; ++k) W.udf[k] = float(W.sc.ud[k]);
    W.critf = float(W.sc.crit);
    tile_ball(g, ix0, iy0, iz0, A.tile, A.cphi, A.sphi, W.tc, W.rho);
    for (int k = 0; k < 3; ++k) W.tc[k] = uniform_d(W.tc[k]);
    W.rho = uniform_d(W.rho);
    tile_ball_f(W.tc, W.rho, W.tcf, W.rhof);
    for (int k = 0; k < 3; ++k) W.tcf[k] = uniform_f(W.tcf[k]);
    W.rhof = uniform_f(W.rhof);
    // plate gates that no ray from any point of bar d (inside its bounding
    // ball, the half diagonal with a rounding margin) to the tile can trigger
    const double qc[3] = {bd.cx, bd.cy, bd.cz};
    const double rb = 0.5 * bd.rdiag * (1.0 + 1e-12) + 1e-9;
    const bool clear = lane < g.n_layers &&
                       plate_gate_clear_ball(G.layers_d[lane < g.n_layers ? lane : 0], qc, rb, W.tc, W.rho);
    W.bar_gate = (uint32_t)__ballot(clear);
}

// Step 1: candidate bars and holes of the tile: those whose box a segment from
// bar d to the tile can meet (bar_tile_may_hit, first on the box of each
// group of kGroup consecutive bars or holes, then on the members of the
// groups it keeps; any other has chord 0 for every ray of the tile, so its
// culling decision does not matter), each with the ballot of the lanes whose
// reference culling test keeps it.  Returns false on overflow.
template <bool kExact, class Lists>
__device__ bool build_candidates(const PeLaunch& A, int d, int lane, const WaveSetup& W,
                                 Lists& L, int& nb, int& nh) {
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const Bar<float>& bdf = G.bars_f[d];
    nb = 0;
    nh = 0;
    for (int w = lane; w < (L.cap_bars + 63) / 64; w += 64) L.tile_far[w] = 0ull;
    // groups of kGroup bars first (one group box per lane), then the members
    // of each group that can be hit
    const int n_bg = (g.n_det + kGroup - 1) / kGroup;
    for (int g0 = 0; g0 < n_bg; g0 += 64) {
        bool gpass = false;
        if (g0 + lane < n_bg) {
            const GroupBox& B = G.bar_groups[g0 + lane];
            gpass = bar_tile_may_hit(bdf, W.tcf, W.rhof, B.c[0], B.c[1], B.c[2], 1.0f, 0.0f, B.h[0],
                                     B.h[1], B.h[2]);
        }
        uint64_t gbal = __ballot(gpass);
        while (gbal) {
            const int base = (g0 + __ffsll((unsigned long long)gbal) - 1) * kGroup;
            gbal &= gbal - 1;
            int k = base + lane;
            bool pass = false, far = false;
            if (k < g.n_det && k != d) {
                const Bar<float>& b = G.bars_f[k];
                pass = bar_tile_may_hit(bdf, W.tcf, W.rhof, b.cx, b.cy, b.cz, b.cb, b.sb, 0.5f * b.w,
                                        0.5f * b.t, 0.5f * b.h);
                if (pass) far = fast_tile_far(G, k, W.tc, W.rho);
            }
            uint64_t bal = __ballot(pass);
            const uint64_t far_bal = __ballot(far);
            while (bal) {
                int bit = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                int kk = base + bit;
                bool keep = W.valid && cull_bar<kExact>(G, kk, W.p, W.pf, W.sc, W.udf, W.critf);
                uint64_t m = __ballot(keep);
                if (m) {
                    if (nb < L.cap_bars && lane == 0) {
                        L.bar_idx[nb] = (uint16_t)kk;
                        L.bar_mask[nb] = m;
                        if ((far_bal >> bit) & 1ull) L.tile_far[nb >> 6] |= 1ull << (nb & 63);
                    }
                    ++nb;
                }
            }
        }
    }
    const int n_hg = (g.n_holes + kGroup - 1) / kGroup;
    for (int g0 = 0; g0 < n_hg; g0 += 64) {
        bool gpass = false;
        if (g0 + lane < n_hg) {
            const GroupBox& B = G.hole_groups[g0 + lane];
            gpass = bar_tile_may_hit(bdf, W.tcf, W.rhof, B.c[0], B.c[1], B.c[2], 1.0f, 0.0f, B.h[0],
                                     B.h[1], B.h[2]);
        }
        uint64_t gbal = __ballot(gpass);
        while (gbal) {
            const int base = (g0 + __ffsll((unsigned long long)gbal) - 1) * kGroup;
            gbal &= gbal - 1;
            int h = base + lane;
            bool pass = false;
            if (h < g.n_holes) {
                const Hole<float>& o = G.holes_f[h];
                pass = bar_tile_may_hit(bdf, W.tcf, W.rhof, o.xc, o.ymid, o.zc, 1.0f, 0.0f, o.r,
                                        0.5f * tabs(o.y2 - o.y1), o.r);
            }
            uint64_t bal = __ballot(pass);
            while (bal) {
                int bit = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                int hh = base + bit;
                bool keep = W.valid && cull_hole<kExact>(G, hh, W.p, W.pf, W.sc, W.udf, W.critf);
                uint64_t m = __ballot(keep);
                if (m) {
                    if (nh < L.cap_holes && lane == 0) {
                        L.hole_idx[nh] = (uint16_t)hh;
                        L.hole_mask[nh] = m;
                    }
                    ++nh;
                }
            }
        }
    }
    return nb <= L.cap_bars && nh <= L.cap_holes;
}

// Sub-cell `cell` of bar d for this wave's tile (from the sub-cell table when
// there is one), with the plate gates that no ray of the tile can trigger:
// those cleared for the whole bar, else tested for the sub-cell point (one
// layer per lane, then a ballot).
__device__ __forceinline__ void tile_cell(const PeLaunch& A, const Bar<double>& bd, int d, int cell,
                                          const WaveSetup& W, int lane, CellCtx& cc) {
    if (A.cells)
        cc = A.cells[(int64_t)(d - A.cells_bar0) * (A.sub[0] * A.sub[1] * A.sub[2]) + cell];
    else
        make_cell_ctx_linear(bd, cell, A.sub, cc);
    for (int k = 0; k < 3; ++k) {
        cc.q[k] = uniform_d(cc.q[k]);
        cc.qs[k] = uniform_d(cc.qs[k]);
        cc.lo[k] = uniform_d(cc.lo[k]);
        cc.hi[k] = uniform_d(cc.hi[k]);
        cc.area[k] = uniform_d(cc.area[k]);
        cc.hc[k] = uniform_f(cc.hc[k]);
    }
    const GeomDesc& g = A.geom.g;
    const uint32_t all = (uint32_t)((1ull << g.n_layers) - 1ull);
    if ((W.bar_gate & all) == all) {  // wave-uniform
        cc.gate_clear = W.bar_gate;
        return;
    }
    const bool clear = lane < g.n_layers && !((W.bar_gate >> lane) & 1u) &&
                       plate_gate_clear(A.geom.layers_d[lane < g.n_layers ? lane : 0], cc.q, W.tc, W.rho);
    cc.gate_clear = W.bar_gate | (uint32_t)__ballot(clear);
}

// Step 2: candidates whose box a segment from the sub-cell centre qd to the
// tile can meet (box_meets_frustum with a point source), bars and holes in
// one pass; writes their list indices to bar_sel / hole_sel.
template <class Lists>
__device__ void narrow_candidates(const PeLaunch& A, const double qd[3], const WaveSetup& W,
                                  int lane, Lists& L, int nb, int nh, int& nsb, int& nsh) {
    const PeGeom& G = A.geom;
    const float q[3] = {float(qd[0]), float(qd[1]), float(qd[2])};
    const float rq[3] = {0.0f, 0.0f, 0.0f};
    nsb = 0;
    nsh = 0;
    const int n = nb + nh;
    for (int base = 0; base < n; base += 64) {
        const int j = base + lane;
        const bool is_bar = j < nb;
        bool pass = false;
        if (j < n) {
            float c[3], cb, sb, h[3];
            if (is_bar) {
                const Bar<float>& bf = G.bars_f[L.bar_idx[j]];
                c[0] = bf.cx; c[1] = bf.cy; c[2] = bf.cz;
                cb = bf.cb; sb = bf.sb;
                h[0] = 0.5f * bf.w; h[1] = 0.5f * bf.t; h[2] = 0.5f * bf.h;
            } else {
                const Hole<float>& o = G.holes_f[L.hole_idx[j - nb]];
                c[0] = o.xc; c[1] = o.ymid; c[2] = o.zc;
                cb = 1.0f; sb = 0.0f;
                h[0] = o.r; h[1] = 0.5f * tabs(o.y2 - o.y1); h[2] = o.r;
            }
            pass = box_meets_frustum(q, rq, W.tcf, W.rhof, c[0], c[1], c[2], cb, sb, h[0], h[1], h[2]);
        }
        const uint64_t bb = __ballot(pass && is_bar), bh = __ballot(pass && !is_bar);
        if (pass) {
            if (is_bar) {
                const int pos = nsb + lanes_below(bb, lane);
                L.bar_sel[pos] = (uint16_t)j;
                if (pos < kFastBars) {
                    float f[6];
                    // fast_faces: the tile part once per tile (build_candidates)
                    const bool near_q_ok = fast_faces_q(G, L.bar_idx[j], qd, f);
                    L.fast_ok[pos] = ((L.tile_far[j >> 6] >> (j & 63)) & 1ull) && near_q_ok ? 1 : 0;
#pragma unroll
                    for (int i = 0; i < 6; ++i) L.faces[pos][i] = f[i];
                    L.faces[pos][6] = G.bars_f[L.bar_idx[j]].mut;
                }
            } else {
                L.hole_sel[nsh + lanes_below(bh, lane)] = (uint16_t)(j - nb);
            }
        }
        nsb += __popcll(bb);
        nsh += __popcll(bh);
    }
}


// Point weights of the crystal grid scatterers (kPoints; an extension with no
// reference counterpart): the sub-cells of bar d are the grid cells, and for
// each the kernel writes the first-leg weight of its centre,
//   dv mu_c,d / (4 pi |q - p|^2) exp(-att),
// att being the PE ray's attenuation up to the point (the sub-cell's own
// chord in bar d added back), to pts[((d - bar0) n_cells + cell) V + v].
// FP64 form (the exact path and FP64 re-computations); the FP32 path forms
// the same product in FP32 (relative error ~1e-7).
__device__ __forceinline__ float point_weight(const Bar<double>& bd, int n_cells, const double aux[8]) {
    const double dv = bd.w * bd.t * bd.h / n_cells;
    return float(aux[5] * exp(-aux[4] * bd.mut) * dv * bd.muc /
                 (4.0 * 3.14159265358979323846 * aux[6]));
}

__device__ __forceinline__ float point_weight_f(const Bar<double>& bd, int n_cells, const double aux[8]) {
    const float c = float(bd.w * bd.t * bd.h / n_cells * bd.muc / (4.0 * 3.14159265358979323846));
    return float(aux[5]) * exp_neg<float>(aux[4] * bd.mut) * tdiv(c, float(aux[6]));
}

// Point weights: true when some row in A.use_rows may use the weight of grid
// point qs for a voxel of the wave's tile.  Lane l tests row l: with the
// tile's voxel centres in the ball (tc, rho) and the row's sub-cell centres in
// its bar's bounding ball (Q, rb), every scattering angle at qs is at least
// angle(qs - tc, Q - qs) - asin(rho / |qs - tc|) - asin(rb / |Q - qs|); when
// its cosine is 1.1e-4 below the row's window cut, the scatter kernel's
// per-lane window test rejects every such term before it reads the weight.
__device__ bool point_used(const PeLaunch& A, const WaveSetup& W, const double qs[3], int lane) {
    bool use = false;
    if (lane < A.n_use_rows) {
        const double* R = A.use_rows + 5 * lane;
        const double c_lo = R[4];
        double u[3], b[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = qs[k] - W.tc[k];
            b[k] = R[k] - qs[k];
        }
        const double L = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double B = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        if (c_lo <= -1.5 || !(L > W.rho * (1.0 + 1e-9) + 1e-9) || !(B > R[3] * (1.0 + 1e-9) + 1e-9)) {
            use = true;
        } else {
            double cu = (u[0] * b[0] + u[1] * b[1] + u[2] * b[2]) / (L * B);
            cu = cu > 1.0 ? 1.0 : (cu < -1.0 ? -1.0 : cu);
            const double th = acos(cu) - asin(W.rho / L) - asin(R[3] / B);
            use = !(th > 0.0) || !(cos(th) < c_lo - 1.1e-4);
        }
    }
    return __ballot(use) != 0;
}

// Work item `item` of the launch (a block index: bar, tile and, for point
// weights, slice of cells) with candidate lists L.  Returns false, having
// written nothing, when the candidates overflow L.
template <bool kExact, bool kPoints, bool kEll, class Lists>
__device__ __forceinline__ bool pe_forward_body(const PeLaunch& A, int item, Lists& L) {
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    // point weights: cells [c_begin, c_end) of the wave's slice; PE: all cells
    const int slices = kPoints ? A.cell_slices : 1;
    const int wave = item / slices;
    const int d = A.bar0 + wave / n_tiles;
    const int tile = wave % n_tiles;
    float* out = A.dst + (int64_t)d * g.n_vox;
    const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
    const int per = (n_cells + slices - 1) / slices;
    const int c_begin = kPoints ? (item % slices) * per : 0;
    const int c_end = kPoints ? min(n_cells, c_begin + per) : n_cells;
    float* pts = kPoints ? A.pts + (int64_t)(d - A.bar0) * n_cells * g.n_vox : nullptr;
    const Bar<double>& bd = G.bars_d[d];
    WaveSetup W;
    wave_setup(A, d, tile, lane, W);
    if (!(bd.mut > (double)0.01f)) {  // PESysMatGen.cu:1116, :1194
        if constexpr (kPoints) {
            if (W.valid)
                for (int c = c_begin; c < c_end; ++c) pts[(int64_t)c * g.n_vox + W.v] = 0.0f;
        } else {
            if (W.valid) out[W.v] = 0.0f;
        }
        return true;
    }
    int nb, nh;
    if (!build_candidates<kExact>(A, d, lane, W, L, nb, nh)) return false;
    __syncthreads();

    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    unsigned fallbacks = 0;
    // Sub-cells in the reference's loop order (z outer, then x, then y;
    // PESysMatGen.cu:1193-1197).  FP32 path: a ray whose decisions came
    // within tolerance of a threshold is not added; its sub-cell goes to a
    // per-wave queue with the ballot of such lanes, and the queue is redone
    // in FP64 when full and at the end.  Keeping the FP64 ray out of the FP32
    // loop keeps the loop's register use low.
    __shared__ uint16_t redo_cell[kRedoCap];
    __shared__ uint64_t redo_mask[kRedoCap];
    int cell = c_begin;
    while (cell < c_end) {
        int n_redo = 0;
        for (; cell < c_end && n_redo < kRedoCap; ++cell) {
            CellCtx cc;
            tile_cell(A, bd, d, cell, W, lane, cc);
            if constexpr (kPoints) {
                if (A.n_use_rows > 0 && !point_used(A, W, cc.q, lane)) {  // wave-uniform
                    if (W.valid) pts[(int64_t)cell * g.n_vox + W.v] = 0.0f;
                    continue;
                }
            }
            int nsb, nsh;
            __syncthreads();
            narrow_candidates(A, cc.q, W, lane, L, nb, nh, nsb, nsh);
            __syncthreads();
            bool near = false;
            if (W.valid) {
                ListCands<Lists> C{&L, nsb, nsh, lane};
                if constexpr (kPoints) {
                    double aux[8];
                    pe_cell_ray<T, kEll>(G, d, W.p, W.ps, cc, C, near, nullptr, aux);
                    if (!near)
                        pts[(int64_t)cell * g.n_vox + W.v] =
                            kExact ? point_weight(bd, n_cells, aux) : point_weight_f(bd, n_cells, aux);
                } else {
                    T val = pe_cell_ray<T, kEll>(G, d, W.p, W.ps, cc, C, near);
                    if (!near) acc.add(val);
                }
            }
            if constexpr (!kExact) {
                const uint64_t m = __ballot(near);
                if (m) {
                    if (lane == 0) {
                        redo_cell[n_redo] = (uint16_t)cell;
                        redo_mask[n_redo] = m;
                    }
                    ++n_redo;
                }
            }
        }
        if constexpr (!kExact) {
            for (int i = 0; i < n_redo; ++i) {
                CellCtx cc;
                tile_cell(A, bd, d, redo_cell[i], W, lane, cc);
                int rsb, rsh;
                __syncthreads();
                narrow_candidates(A, cc.q, W, lane, L, nb, nh, rsb, rsh);
                __syncthreads();
                if ((redo_mask[i] >> lane) & 1ull) {
                    ListCands<Lists> C{&L, rsb, rsh, lane};
                    bool n2 = false;
                    if constexpr (kPoints) {
                        double aux[8];
                        pe_cell_ray<double, kEll>(G, d, W.p, W.ps, cc, C, n2, nullptr, aux);
                        pts[(int64_t)redo_cell[i] * g.n_vox + W.v] = point_weight(bd, n_cells, aux);
                    } else {
                        acc.add((T)pe_cell_ray<double, kEll>(G, d, W.p, W.ps, cc, C, n2));
                    }
                    ++fallbacks;
                }
            }
        }
    }
    if constexpr (!kPoints) {
        if (W.valid) out[W.v] = float(acc.value());
    }
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
    return true;
}

// 4 waves per SIMD: 128 VGPRs (the sub-cell context and the tile ball in
// scalar registers) with a ~20-byte spill, and 16 one-wave blocks per CU
// within the LDS (10 KB of lists each).  The projections below likewise.
// kEll: elliptical holes (the elliptic-cylinder chord compiled in; the
// all-FP64 kernels always have it).
template <bool kExact, bool kEll = kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void pe_forward_kernel(PeLaunch A) {
    __shared__ WaveLists L;
    const int item = xcd_item((int)blockIdx.x, (int)gridDim.x);
    if (!pe_forward_body<kExact, false, kEll>(A, item, L)) spill_push(A, item, threadIdx.x);
}

// Point weights at 3 waves per SIMD (the register cap costs no spills here;
// without it the allocation lands one register above the 3-wave limit).
template <bool kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void pe_points_kernel(PeLaunch A) {
    __shared__ WaveLists L;
    const int item = xcd_item((int)blockIdx.x, (int)gridDim.x);
    if (!pe_forward_body<kExact, true, kExact>(A, item, L)) spill_push(A, item, threadIdx.x);
}

// The work items the kernel above queued (counters[2] of them), redone with
// candidate lists in global memory: wave b takes queue entries b, b + grid, ...
// and its own list slot.  The entry count is wave-uniform, so every wave ends.
template <bool kExact, bool kPoints, bool kEll = kExact>
__global__ __launch_bounds__(64) void pe_forward_spill_kernel(PeLaunch A) {
    SpillLists L = spill_lists(A.spill, (int)blockIdx.x);
    const unsigned long long n = A.counters[2];
    for (unsigned long long i = blockIdx.x; i < n; i += gridDim.x) {
        __syncthreads();  // the previous item's list reads are done
        pe_forward_body<kExact, kPoints, kEll>(A, A.spill.queue[i], L);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&A.counters[1], n);
}

// ---- tangent --------------------------------------------------------------------------
//
// d term of one sub-cell ray along the direction v (no counterpart in the
// reference; the derivatives of its element formula, PESysMatGen.cu:1240-1357,
// as in the adjoint): term = Omega e^-att abs,
//   d att = sum_m v_m rest_m + sum_k v_t[k] l_k + v_t[d] (l1 - l2),
//   d abs = v_t[d] mu_pe (l_a e^(-mu l_a)/mu - (1 - e^(-mu l_a))/mu^2) + v_pe[d] (1 - e^(-mu l_a))/mu,
//   d term = -term d att + Omega e^-att d abs.
// dlay, aux and tan_att as pe_cell_ray returns them.
__device__ __forceinline__ double tangent_term(const PeGeom& G, int d, const double* v, double term,
                                               const double* dlay, const double* aux, double tan_att) {
    const GeomDesc& g = G.g;
    const Bar<double>& bd = G.bars_d[d];
    const double vt = v[g.n_layers + d], vpe = v[g.n_layers + g.n_det + d];
    double datt = tan_att + vt * aux[1];
#pragma unroll
    for (int m = 0; m < kMaxLayers; ++m)
        if (m < g.n_layers) datt += v[m] * dlay[m];
    const double mut = bd.mut, la = aux[2], ea = aux[3];
    const double dabs = vt * bd.mupe * (la * ea / mut - (1.0 - ea) / (mut * mut)) + vpe * (1.0 - ea) / mut;
    return -term * datt + aux[0] * dabs;
}

// Work item `item` (bar, tile) of the tangent: as pe_forward_body, with the
// derivative summed beside the element.  Returns false, having written
// nothing, when the candidates overflow L.
template <bool kExact, class Lists>
__device__ __forceinline__ bool pe_tangent_body(const PeLaunch& A, const PeTangent& TG, int item, Lists& L) {
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int d = item / n_tiles;
    const int tile = item % n_tiles;
    const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
    const Bar<double>& bd = G.bars_d[d];
    WaveSetup W;
    wave_setup(A, d, tile, lane, W);
    const int64_t o = (int64_t)d * g.n_vox + W.v;
    if (!(bd.mut > (double)0.01f)) {  // PESysMatGen.cu:1116, :1194
        if (W.valid) {
            TG.d_out[o] = 0.0f;
            if (TG.a_out) TG.a_out[o] = 0.0f;
        }
        return true;
    }
    int nb, nh;
    if (!build_candidates<kExact>(A, d, lane, W, L, nb, nh)) return false;
    __syncthreads();
    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    KahanSum<double> dacc;
    unsigned fallbacks = 0;
    const double* vbar = TG.v + g.n_layers;
    __shared__ uint16_t redo_cell[kRedoCap];
    __shared__ uint64_t redo_mask[kRedoCap];
    int cell = 0;
    while (cell < n_cells) {
        int n_redo = 0;
        for (; cell < n_cells && n_redo < kRedoCap; ++cell) {
            CellCtx cc;
            tile_cell(A, bd, d, cell, W, lane, cc);
            int nsb, nsh;
            __syncthreads();
            narrow_candidates(A, cc.q, W, lane, L, nb, nh, nsb, nsh);
            __syncthreads();
            bool near = false;
            if (W.valid) {
                ListCands<Lists> C{&L, nsb, nsh, lane};
                double dlay[kMaxLayers], aux[8], tatt = 0.0;
                T val = pe_cell_ray<T, kExact>(G, d, W.p, W.ps, cc, C, near, dlay, aux, nullptr, 0, vbar, &tatt);
                if (!near) {
                    acc.add(val);
                    dacc.add(tangent_term(G, d, TG.v, (double)val, dlay, aux, tatt));
                }
            }
            if constexpr (!kExact) {
                const uint64_t m = __ballot(near);
                if (m) {
                    if (lane == 0) {
                        redo_cell[n_redo] = (uint16_t)cell;
                        redo_mask[n_redo] = m;
                    }
                    ++n_redo;
                }
            }
        }
        if constexpr (!kExact) {
            for (int i = 0; i < n_redo; ++i) {
                CellCtx cc;
                tile_cell(A, bd, d, redo_cell[i], W, lane, cc);
                int rsb, rsh;
                __syncthreads();
                narrow_candidates(A, cc.q, W, lane, L, nb, nh, rsb, rsh);
                __syncthreads();
                if ((redo_mask[i] >> lane) & 1ull) {
                    ListCands<Lists> C{&L, rsb, rsh, lane};
                    bool n2 = false;
                    double dlay[kMaxLayers], aux[8], tatt = 0.0;
                    const double val = pe_cell_ray<double, kExact>(G, d, W.p, W.ps, cc, C, n2, dlay, aux,
                                                                   nullptr, 0, vbar, &tatt);
                    acc.add((T)val);
                    dacc.add(tangent_term(G, d, TG.v, val, dlay, aux, tatt));
                    ++fallbacks;
                }
            }
        }
    }
    if (W.valid) {
        TG.d_out[o] = float(dacc.value());
        if (TG.a_out) TG.a_out[o] = float(acc.value());
    }
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
    return true;
}

template <bool kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void pe_tangent_kernel(PeLaunch A, PeTangent TG) {
    __shared__ WaveLists L;
    if (!pe_tangent_body<kExact>(A, TG, (int)blockIdx.x, L)) spill_push(A, (int)blockIdx.x, threadIdx.x);
}

template <bool kExact>
__global__ __launch_bounds__(64) void pe_tangent_spill_kernel(PeLaunch A, PeTangent TG) {
    SpillLists L = spill_lists(A.spill, (int)blockIdx.x);
    const unsigned long long n = A.counters[2];
    for (unsigned long long i = blockIdx.x; i < n; i += gridDim.x) {
        __syncthreads();
        pe_tangent_body<kExact>(A, TG, A.spill.queue[i], L);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&A.counters[1], n);
}

// ---- adjoint -------------------------------------------------------------------------
//
// dL/dmu of L = sum w[d][v] A[d][v].  Per sub-cell ray with term c (the
// forward contribution) and weight w (from the derivatives of the reference
// element formula, PESysMatGen.cu:1240-1357):
//   layer m:         -w c (T_m / cos(alpha) - sum of hole chords), when counted
//   bar k != d:      -w c l_k
//   mu_total of d:   -w c (l1 - l2) + w Omega e^-att mu_pe (l_a e^(-mu l_a)/mu - (1 - e^(-mu l_a))/mu^2)
//   mu_pe of d:      w Omega e^-att (1 - e^(-mu l_a)) / mu
// The chords through other bars are recomputed after the forward term, not
// stored.  One wave handles detector d and work items g, g + G, ... (an item
// is a tile, or a tile's range of sub-cells when bars x tiles alone would
// leave the GPU under-filled); partial sums
// go to a per-(d, g) row of the scratch buffer (bar mu_total partials dense
// over all bars, then the layer, mu_pe and loss slots), added in a fixed
// order, and pe_grad_reduce sums the rows in a fixed order: the result does
// not depend on scheduling.

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// Per-wave adjoint accumulators: per-lane sums for bar d's own mu and the
// loss, wave sums (in LDS, added by lane 0 in a fixed order) for the layers
// and, per tile, for each candidate bar.
struct AdjAcc {
    double own_mut = 0.0, own_mupe = 0.0, loss = 0.0;
};

// Adjoint contributions of one sub-cell for the lanes with `active` set.  In
// FP32 a lane whose ray came within tolerance of a threshold contributes
// nothing and is reported through the return value (the caller redoes it in
// FP64).  Every lane of the wave takes part in the wave sums.
template <class T, bool kEll, class Lists>
__device__ bool adjoint_cell(const PeGeom& G, int d, int lane, const WaveSetup& W,
                             const Lists& L, const ListCands<Lists>& C, const CellCtx& cc,
                             bool active, double wv, AdjAcc& acc, double* lay_acc,
                             double* cand_acc, float* keep = nullptr) {
    const GeomDesc& g = G.g;
    const Bar<double>& bd = G.bars_d[d];
    const double mut = bd.mut, mupe = bd.mupe;
    double term = 0.0, dlay[kMaxLayers], aux[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < kMaxLayers; ++m) dlay[m] = 0.0;
    bool near = false;
    if (active) {
        term = (double)pe_cell_ray<T, kEll>(G, d, W.p, W.ps, cc, C, near, dlay, aux, keep, 64);
        if (kGuarded<T> && near) active = false;
    }
    const double cw = active ? wv * term : 0.0;
    if (active) {
        const double la = aux[2], ea = aux[3], effom = aux[0];
        acc.loss += cw;
        acc.own_mut += -cw * aux[1] + wv * effom * mupe * (la * ea / mut - (1.0 - ea) / (mut * mut));
        acc.own_mupe += wv * effom * (1.0 - ea) / mut;
    }
#pragma unroll
    for (int m = 0; m < kMaxLayers; ++m) {
        if (m < g.n_layers) {
            const double x = wave_sum(-cw * dlay[m]);
            if (lane == 0) lay_acc[m] += x;
        }
    }
    // chords through the other bars: kept from the ray (the first kKeepChords
    // candidates, FP32 rays) or recomputed
    double D[3], dd = 0.0, inv_dd = 0.0;
    SegDir<T> sw;
    T tol_d = T(0);
    if (!keep || C.nbars() > kKeepChords) {  // wave-uniform
        bool nd = false;
        ray_dir<T>(W.p, cc.q, D, dd, sw, tol_d, nd);
        inv_dd = 1.0 / dd;
    }
    for (int j = 0; j < C.nbars(); ++j) {
        double x = 0.0;
        if (active && cw != 0.0 && C.bar_on(j)) {
            double l;
            if (keep && j < kKeepChords) {
                l = (double)keep[j * 64];
            } else {
                bool n3 = false;
                l = (double)other_bar_chord<T>(G, C.bar(j), W.p, D, dd, inv_dd, cc.q, sw, tol_d, n3);
            }
            x = -cw * l;
        }
        x = wave_sum(x);
        if (lane == 0) cand_acc[L.bar_sel[j]] += x;
    }
    return near;
}

// Adjoint of work item `item` (a block index: bar d, group of tiles) with
// candidate lists L and per-candidate sums cand_acc.  Returns false when the
// candidates overflow L (its scratch row is then incomplete; the spill kernel
// rewrites all of it).
template <bool kExact, class Lists>
__device__ __forceinline__ bool pe_backward_body(const PeLaunch& A, const PeGradScratch& S, int item,
                                                 Lists& L, double* cand_acc) {
    __shared__ double lay_acc[kMaxLayers];
    __shared__ float keep[kKeepChords * 64];  // FP32 rays' chords, [candidate][lane]
    __shared__ uint16_t redo_cell[kRedoCap];
    __shared__ uint64_t redo_mask[kRedoCap];
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int d = item / S.groups;
    const int grp = item % S.groups;
    const Bar<double>& bd = G.bars_d[d];
    double* out_row = S.rows + (int64_t)item * S.stride;
    // the bars' mu_total partials accumulate in this wave's own row of the
    // scratch buffer (no other wave writes it; LDS would cap the occupancy)
    for (int k = lane; k < g.n_det; k += 64) out_row[k] = 0.0;
    if (lane < kMaxLayers) lay_acc[lane] = 0.0;
    AdjAcc acc;
    unsigned fallbacks = 0;
    bool overflow = false;
    __syncthreads();
    if (bd.mut > (double)0.01f) {
        using T = typename std::conditional<kExact, double, float>::type;
        const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
        // work items: (tile, range of sub-cells); item i is tile i / cs, range i % cs
        const int cs = S.cell_slices;
        const int per = (n_cells + cs - 1) / cs;
        for (int item = grp; item < n_tiles * cs; item += S.groups) {
            const int tile = item / cs;
            const int c_begin = (item % cs) * per, c_end = min(n_cells, c_begin + per);
            WaveSetup W;
            wave_setup(A, d, tile, lane, W);
            int nb, nh;
            if (!build_candidates<kExact>(A, d, lane, W, L, nb, nh)) {
                overflow = true;
                break;
            }
            for (int j = lane; j < nb; j += 64) cand_acc[j] = 0.0;
            const double wv = W.valid ? (double)S.w[(int64_t)d * g.n_vox + W.v] : 0.0;
            __syncthreads();
            // sub-cells in the reference's order; FP32 rays near a threshold
            // are queued and redone in FP64 (as in the forward kernel)
            int cell = c_begin;
            while (cell < c_end) {
                int n_redo = 0;
                for (; cell < c_end && n_redo < kRedoCap; ++cell) {
                    CellCtx cc;
                    tile_cell(A, bd, d, cell, W, lane, cc);
                    int nsb, nsh;
                    __syncthreads();
                    narrow_candidates(A, cc.q, W, lane, L, nb, nh, nsb, nsh);
                    __syncthreads();
                    ListCands<Lists> C{&L, nsb, nsh, lane};
                    const bool near = adjoint_cell<T, kExact>(G, d, lane, W, L, C, cc, W.valid, wv, acc,
                                                      lay_acc, cand_acc, kExact ? nullptr : keep + lane);
                    if constexpr (!kExact) {
                        const uint64_t m = __ballot(near);
                        if (m) {
                            if (lane == 0) {
                                redo_cell[n_redo] = (uint16_t)cell;
                                redo_mask[n_redo] = m;
                            }
                            ++n_redo;
                        }
                    }
                    __syncthreads();
                }
                if constexpr (!kExact) {
                    for (int i = 0; i < n_redo; ++i) {
                        CellCtx cc;
                        tile_cell(A, bd, d, redo_cell[i], W, lane, cc);
                        int rsb, rsh;
                        __syncthreads();
                        narrow_candidates(A, cc.q, W, lane, L, nb, nh, rsb, rsh);
                        __syncthreads();
                        ListCands<Lists> C{&L, rsb, rsh, lane};
                        const bool mine = (redo_mask[i] >> lane) & 1ull;
                        adjoint_cell<double, kExact>(G, d, lane, W, L, C, cc, mine, wv, acc, lay_acc, cand_acc);
                        fallbacks += mine;
                    }
                }
            }
            // fold this tile's candidate partials into the dense row (one
            // entry per candidate, each bar at most once per tile)
            for (int j = lane; j < nb; j += 64) out_row[L.bar_idx[j]] += cand_acc[j];
            __syncthreads();
        }
    }
    const double own_mut = wave_sum(acc.own_mut);
    const double own_mupe = wave_sum(acc.own_mupe);
    const double loss = wave_sum(acc.loss);
    __syncthreads();  // the lanes' row updates before lane 0 adds to out_row[d]
    if (lane == 0) {
        out_row[d] += own_mut;
        for (int m = 0; m < g.n_layers; ++m) out_row[g.n_det + m] = lay_acc[m];
        out_row[g.n_det + g.n_layers] = own_mupe;
        out_row[g.n_det + g.n_layers + 1] = loss;
    }
    if (overflow) return false;
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
    return true;
}

// 3 waves per SIMD (LDS and a 168-VGPR cap); uncapped the allocation lands
// at 188 VGPRs, 2 waves.
template <bool kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void pe_backward_kernel(
    PeLaunch A, PeGradScratch S) {
    __shared__ AdjointLists L;
    __shared__ double cand_acc[kCapBarsAdjoint];
    if (!pe_backward_body<kExact>(A, S, (int)blockIdx.x, L, cand_acc)) spill_push(A, (int)blockIdx.x, threadIdx.x);
}

// Queued work items of pe_backward_kernel redone with lists in global memory
// (as pe_forward_spill_kernel); each rewrites its whole scratch row.
template <bool kExact>
__global__ __launch_bounds__(64) void pe_backward_spill_kernel(PeLaunch A, PeGradScratch S) {
    SpillLists L = spill_lists(A.spill, (int)blockIdx.x);
    const unsigned long long n = A.counters[2];
    for (unsigned long long i = blockIdx.x; i < n; i += gridDim.x) {
        __syncthreads();
        pe_backward_body<kExact>(A, S, A.spill.queue[i], L, L.cand_acc);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&A.counters[1], n);
}

// grad[0 .. Lc-1]: layers; grad[Lc + k]: mu_total of bar k; grad[Lc + N + k]:
// mu_pe of bar k; grad[Lc + 2N]: the loss.  One block per entry, rows in a fixed order.
__global__ __launch_bounds__(256) void pe_grad_reduce(PeGradScratch S, int n_det, int n_layers,
                                                      double* grad) {
    // one block per entry: thread t sums rows t, t + 256, ... in order, then a
    // fixed tree over the threads (the same order on every run)
    __shared__ double part[256];
    const int i = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t n_rows = (int64_t)n_det * S.groups;
    double s = 0.0;
    if (i < n_layers) {
        for (int64_t r = t; r < n_rows; r += 256) s += S.rows[r * S.stride + n_det + i];
    } else if (i < n_layers + n_det) {
        const int k = i - n_layers;
        for (int64_t r = t; r < n_rows; r += 256) s += S.rows[r * S.stride + k];
    } else if (i < n_layers + 2 * n_det) {
        const int k = i - n_layers - n_det;
        for (int grp = t; grp < S.groups; grp += 256)
            s += S.rows[((int64_t)k * S.groups + grp) * S.stride + n_det + n_layers];
    } else {
        for (int64_t r = t; r < n_rows; r += 256) s += S.rows[r * S.stride + n_det + n_layers + 1];
    }
    part[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) part[t] += part[t + h];
        __syncthreads();
    }
    if (t == 0) grad[i] = part[0];
}

// ---- matrix-free projections -----------------------------------------------------------
//
// (No counterpart in the reference, which writes the matrix.)  The element
// A[d][v] of one lane, formed as in pe_forward_kernel (same sub-cell order,
// same FP64 redo of near rays), is used at once instead of being stored:
// forward projection y[d] = sum_v A[d][v] x[v], back-projection
// x[v] = sum_d A[d][v] y[d].  Sums are FP64 in a fixed order.

// A[d][v] of this lane for bar d and the wave's tile (W set up for them) into
// a: 0 for a bar below the mu threshold.  Returns false (wave-uniform) when
// the candidates overflow L; adds the lane's FP64 re-computations to fallbacks.
template <bool kExact, bool kEll, class Lists>
__device__ bool pe_element(const PeLaunch& A, int d, int lane, const WaveSetup& W, Lists& L,
                           uint16_t* redo_cell, uint64_t* redo_mask, unsigned& fallbacks, double& a) {
    const PeGeom& G = A.geom;
    const Bar<double>& bd = G.bars_d[d];
    a = 0.0;
    if (!(bd.mut > (double)0.01f)) return true;  // PESysMatGen.cu:1116, :1194
    int nb, nh;
    __syncthreads();  // the lists of the previous element are no longer read
    if (!build_candidates<kExact>(A, d, lane, W, L, nb, nh)) return false;
    __syncthreads();
    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
    int cell = 0;
    while (cell < n_cells) {
        int n_redo = 0;
        for (; cell < n_cells && n_redo < kRedoCap; ++cell) {
            CellCtx cc;
            tile_cell(A, bd, d, cell, W, lane, cc);
            int nsb, nsh;
            __syncthreads();
            narrow_candidates(A, cc.q, W, lane, L, nb, nh, nsb, nsh);
            __syncthreads();
            bool near = false;
            if (W.valid) {
                ListCands<Lists> C{&L, nsb, nsh, lane};
                T val = pe_cell_ray<T, kEll>(G, d, W.p, W.ps, cc, C, near);
                if (!near) acc.add(val);
            }
            if constexpr (!kExact) {
                const uint64_t m = __ballot(near);
                if (m) {
                    if (lane == 0) {
                        redo_cell[n_redo] = (uint16_t)cell;
                        redo_mask[n_redo] = m;
                    }
                    ++n_redo;
                }
            }
        }
        if constexpr (!kExact) {
            for (int i = 0; i < n_redo; ++i) {
                CellCtx cc;
                tile_cell(A, bd, d, redo_cell[i], W, lane, cc);
                int rsb, rsh;
                __syncthreads();
                narrow_candidates(A, cc.q, W, lane, L, nb, nh, rsb, rsh);
                __syncthreads();
                if ((redo_mask[i] >> lane) & 1ull) {
                    ListCands<Lists> C{&L, rsb, rsh, lane};
                    bool n2 = false;
                    acc.add((T)pe_cell_ray<double, kEll>(G, d, W.p, W.ps, cc, C, n2));
                    ++fallbacks;
                }
            }
        }
    }
    // the element as the matrix kernel stores it (FP32), so that projecting
    // equals multiplying by the stored matrix
    a = (double)float(acc.value());
    return true;
}

__device__ __forceinline__ void add_fallbacks(const PeLaunch& A, int lane, unsigned fallbacks) {
    unsigned tot = fallbacks;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
}

// Forward projection of work item `item` (bar d, voxel tile): the wave's share
// of y[d] goes to partial[item].  Returns false, having written nothing, when
// the candidates overflow L.
template <bool kExact, bool kEll, class Lists>
__device__ __forceinline__ bool pe_project_body(const PeLaunch& A, const float* x, double* partial, int item,
                                                Lists& L) {
    __shared__ uint16_t redo_cell[kRedoCap];
    __shared__ uint64_t redo_mask[kRedoCap];
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int d = item / n_tiles;
    const int tile = item % n_tiles;
    WaveSetup W;
    wave_setup(A, d, tile, lane, W);
    unsigned fallbacks = 0;
    double a;
    if (!pe_element<kExact, kEll>(A, d, lane, W, L, redo_cell, redo_mask, fallbacks, a)) return false;
    const double s = wave_sum(W.valid ? a * (double)x[W.v] : 0.0);
    if (lane == 0) partial[item] = s;
    if constexpr (!kExact) add_fallbacks(A, lane, fallbacks);
    return true;
}

template <bool kExact, bool kEll = kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void pe_project_kernel(PeLaunch A, const float* x, double* partial) {
    __shared__ WaveLists L;
    if (!pe_project_body<kExact, kEll>(A, x, partial, (int)blockIdx.x, L)) spill_push(A, (int)blockIdx.x, threadIdx.x);
}

// Back-projection of work item `item` (group of `group` consecutive bars,
// voxel tile), looping over the group's bars (those with y[d] = 0 skipped);
// each lane's share of x[v] goes to partial[group][v].  Returns false, having
// written nothing, when the candidates of one of the bars overflow L.
template <bool kExact, bool kEll, class Lists>
__device__ __forceinline__ bool pe_backproject_body(const PeLaunch& A, const float* y, int group,
                                                    double* partial, int item, Lists& L) {
    __shared__ uint16_t redo_cell[kRedoCap];
    __shared__ uint64_t redo_mask[kRedoCap];
    const GeomDesc& g = A.geom.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int grp = item / n_tiles;
    const int tile = item % n_tiles;
    const int d0 = grp * group, d1 = min(g.n_det, d0 + group);
    unsigned fallbacks = 0;
    double acc = 0.0;
    for (int d = d0; d < d1; ++d) {
        const float yd = y[d];
        if (yd == 0.0f) continue;  // wave-uniform
        WaveSetup W;
        wave_setup(A, d, tile, lane, W);
        double a;
        if (!pe_element<kExact, kEll>(A, d, lane, W, L, redo_cell, redo_mask, fallbacks, a)) return false;
        acc += a * (double)yd;
    }
    // the lane's voxel (it does not depend on the bar)
    const int t0 = tile % A.tiles[0], t1 = (tile / A.tiles[0]) % A.tiles[1],
              t2 = tile / (A.tiles[0] * A.tiles[1]);
    const int ix = t0 * A.tile[0] + lane % A.tile[0];
    const int iy = t1 * A.tile[1] + (lane / A.tile[0]) % A.tile[1];
    const int iz = t2 * A.tile[2] + lane / (A.tile[0] * A.tile[1]);
    if (ix < g.nx && iy < g.ny && iz < g.nz)
        partial[(int64_t)grp * g.n_vox + ix + (int64_t)iy * g.nx + (int64_t)iz * g.nx * g.ny] = acc;
    if constexpr (!kExact) add_fallbacks(A, lane, fallbacks);
    return true;
}

template <bool kExact, bool kEll = kExact>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void pe_backproject_kernel(
    PeLaunch A, const float* y, int group, double* partial) {
    __shared__ WaveLists L;
    if (!pe_backproject_body<kExact, kEll>(A, y, group, partial, (int)blockIdx.x, L)) spill_push(A, (int)blockIdx.x, threadIdx.x);
}

// Queued work items of the projection kernels redone with lists in global
// memory (as pe_forward_spill_kernel).
template <bool kExact, bool kBack, bool kEll = kExact>
__global__ __launch_bounds__(64) void pe_project_spill_kernel(PeLaunch A, const float* in, int group,
                                                              double* partial) {
    SpillLists L = spill_lists(A.spill, (int)blockIdx.x);
    const unsigned long long n = A.counters[2];
    for (unsigned long long i = blockIdx.x; i < n; i += gridDim.x) {
        __syncthreads();
        if constexpr (kBack)
            pe_backproject_body<kExact, kEll>(A, in, group, partial, A.spill.queue[i], L);
        else
            pe_project_body<kExact, kEll>(A, in, partial, A.spill.queue[i], L);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&A.counters[1], n);
}

// out[i] = sum_k partial[k * stride + i * step] for k < n, in k order.
__global__ __launch_bounds__(256) void sum_rows_kernel(const double* partial, int64_t n_out, int n,
                                                       int64_t stride, int64_t step, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    KahanSum<double> s;
    for (int k = 0; k < n; ++k) s.add(partial[(int64_t)k * stride + i * step]);
    out[i] = float(s.value());
}

}  // namespace

namespace {

// Checks the spill area against a launch of `blocks` blocks, clears the
// launch's queue length, and returns the spill kernel's grid.  A launch with
// no spill area (queue == nullptr; the caller has cleared counters[2]) gets
// grid 0: its overflowing work items are only counted in counters[2].
hipError_t spill_begin(const PeLaunch& A, int64_t blocks, hipStream_t stream, unsigned& spill_grid) {
    const PeSpill& S = A.spill;
    spill_grid = 0;
    if (!S.queue) return hipSuccess;
    if (!S.slots || S.n_slots < 1 || blocks > S.queue_cap || S.cap_bars < A.geom.g.n_det ||
        S.cap_holes < A.geom.g.n_holes || S.slot_bytes < pe_spill_slot_bytes(S.cap_bars, S.cap_holes))
        return hipErrorInvalidValue;
    spill_grid = (unsigned)std::min<int64_t>(S.n_slots, blocks);
    return hipMemsetAsync(A.counters + 2, 0, sizeof(unsigned long long), stream);
}

}  // namespace

hipError_t launch_pe_tangent(const PeLaunch& A, const PeTangent& TG, bool exact, hipStream_t stream) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int64_t blocks = (int64_t)n_tiles * A.geom.g.n_det;
    if (blocks <= 0) return hipSuccess;
    unsigned sg = 0;
    hipError_t e = spill_begin(A, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_tangent_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A, TG);
        if (sg) hipLaunchKernelGGL(pe_tangent_spill_kernel<true>, dim3(sg), dim3(64), 0, stream, A, TG);
    } else {
        hipLaunchKernelGGL(pe_tangent_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A, TG);
        if (sg) hipLaunchKernelGGL(pe_tangent_spill_kernel<false>, dim3(sg), dim3(64), 0, stream, A, TG);
    }
    return hipGetLastError();
}

hipError_t launch_pe_backward(const PeLaunch& A, const PeGradScratch& S, bool exact, double* grad,
                              hipStream_t stream) {
    const int64_t blocks = (int64_t)A.geom.g.n_det * S.groups;
    if (blocks <= 0) return hipSuccess;
    unsigned sg = 0;
    hipError_t e = spill_begin(A, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_backward_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A, S);
        if (sg) hipLaunchKernelGGL(pe_backward_spill_kernel<true>, dim3(sg), dim3(64), 0, stream, A, S);
    } else {
        hipLaunchKernelGGL(pe_backward_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A, S);
        if (sg) hipLaunchKernelGGL(pe_backward_spill_kernel<false>, dim3(sg), dim3(64), 0, stream, A, S);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n_out = A.geom.g.n_layers + 2 * A.geom.g.n_det + 1;
    hipLaunchKernelGGL(pe_grad_reduce, dim3(n_out), dim3(256), 0, stream, S, A.geom.g.n_det,
                       A.geom.g.n_layers, grad);
    return hipGetLastError();
}

namespace {

// One thread per (bar, sub-cell): the sub-cell context the PE kernels would
// form with make_cell_ctx_linear (same code, same values).
__global__ __launch_bounds__(256) void cell_table_kernel(PeGeom G, int bar0, int n_bars, int s0, int s1,
                                                         int s2, CellCtx* out) {
    const int n_cells = s0 * s1 * s2;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)n_bars * n_cells) return;
    const int b = (int)(tid / n_cells), cell = (int)(tid % n_cells);
    const int sub[3] = {s0, s1, s2};
    CellCtx cc;
    make_cell_ctx_linear(G.bars_d[bar0 + b], cell, sub, cc);
    cc.gate_clear = 0;
    out[tid] = cc;
}

}  // namespace

hipError_t launch_cell_table(const PeGeom& geom, int bar0, int n_bars, const int sub[3], CellCtx* out,
                             hipStream_t stream) {
    const int64_t total = (int64_t)n_bars * sub[0] * sub[1] * sub[2];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(cell_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       geom, bar0, n_bars, sub[0], sub[1], sub[2], out);
    return hipGetLastError();
}

hipError_t launch_pe_forward(const PeLaunch& A, bool exact, hipStream_t stream) {
    return launch_pe_forward_rows(A, exact, 0, A.geom.g.n_det, stream);
}

hipError_t launch_pe_forward_rows(const PeLaunch& A, bool exact, int bar0, int n_bars, hipStream_t stream) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    if (bar0 < 0 || n_bars < 0 || bar0 + n_bars > A.geom.g.n_det) return hipErrorInvalidValue;
    const int64_t blocks = (int64_t)n_tiles * n_bars;
    if (blocks <= 0) return hipSuccess;
    PeLaunch B = A;
    B.bar0 = bar0;
    B.cell_slices = 1;
    unsigned sg = 0;
    hipError_t e = spill_begin(B, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_forward_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, B);
        if (sg) hipLaunchKernelGGL((pe_forward_spill_kernel<true, false>), dim3(sg), dim3(64), 0, stream, B);
    } else if (A.ell) {
        hipLaunchKernelGGL((pe_forward_kernel<false, true>), dim3((unsigned)blocks), dim3(64), 0, stream, B);
        if (sg)
            hipLaunchKernelGGL((pe_forward_spill_kernel<false, false, true>), dim3(sg), dim3(64), 0, stream, B);
    } else {
        hipLaunchKernelGGL(pe_forward_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, B);
        if (sg) hipLaunchKernelGGL((pe_forward_spill_kernel<false, false>), dim3(sg), dim3(64), 0, stream, B);
    }
    return hipGetLastError();
}

hipError_t launch_pe_points(const PeLaunch& A, bool exact, int bar0, int n_bars, hipStream_t stream) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
    const int64_t waves = (int64_t)n_tiles * n_bars;
    if (waves <= 0) return hipSuccess;
    PeLaunch B = A;
    B.bar0 = bar0;
    // cells split over waves until the launch holds about 32768 waves, with
    // at least 64 cells per wave (each wave builds its own candidate lists)
    int64_t slices = (32768 + waves - 1) / waves;
    slices = std::min<int64_t>(slices, std::max(1, n_cells / 64));
    B.cell_slices = (int)std::max<int64_t>(slices, 1);
    const int64_t blocks = waves * B.cell_slices;
    unsigned sg = 0;
    hipError_t e = spill_begin(B, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_points_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, B);
        if (sg) hipLaunchKernelGGL((pe_forward_spill_kernel<true, true>), dim3(sg), dim3(64), 0, stream, B);
    } else {
        hipLaunchKernelGGL(pe_points_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, B);
        if (sg) hipLaunchKernelGGL((pe_forward_spill_kernel<false, true>), dim3(sg), dim3(64), 0, stream, B);
    }
    return hipGetLastError();
}

int backproject_group(const PeLaunch& A) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int64_t waves = (int64_t)A.geom.g.n_det * n_tiles;
    // enough waves to fill the GPU (~32768), at most 64 bars per wave
    return (int)std::min<int64_t>(64, std::max<int64_t>(1, waves / 32768));
}

size_t project_scratch(const PeLaunch& A, bool back) {
    const GeomDesc& g = A.geom.g;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    if (!back) return (size_t)g.n_det * (size_t)n_tiles;
    const int grp = backproject_group(A);
    return (size_t)((g.n_det + grp - 1) / grp) * (size_t)g.n_vox;
}

hipError_t launch_pe_project(const PeLaunch& A, bool exact, const float* x, double* partial, float* y,
                             hipStream_t stream) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int64_t blocks = (int64_t)A.geom.g.n_det * n_tiles;
    if (blocks <= 0) return hipSuccess;
    unsigned sg = 0;
    hipError_t e = spill_begin(A, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_project_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A, x, partial);
        if (sg) hipLaunchKernelGGL((pe_project_spill_kernel<true, false>), dim3(sg), dim3(64), 0, stream, A, x, 0,
                           partial);
    } else if (A.ell) {
        hipLaunchKernelGGL((pe_project_kernel<false, true>), dim3((unsigned)blocks), dim3(64), 0, stream, A, x,
                           partial);
        if (sg)
            hipLaunchKernelGGL((pe_project_spill_kernel<false, false, true>), dim3(sg), dim3(64), 0, stream, A, x,
                               0, partial);
    } else {
        hipLaunchKernelGGL(pe_project_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A, x, partial);
        if (sg) hipLaunchKernelGGL((pe_project_spill_kernel<false, false>), dim3(sg), dim3(64), 0, stream, A, x, 0,
                           partial);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = A.geom.g.n_det;  // y[d] = sum over tiles of partial[d * n_tiles + t]
    hipLaunchKernelGGL(sum_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, partial,
                       n, n_tiles, (int64_t)1, (int64_t)n_tiles, y);
    return hipGetLastError();
}

hipError_t launch_pe_backproject(const PeLaunch& A, bool exact, const float* y, double* partial, float* x,
                                 hipStream_t stream) {
    const GeomDesc& g = A.geom.g;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int grp = backproject_group(A);
    const int n_grp = (g.n_det + grp - 1) / grp;
    const int64_t blocks = (int64_t)n_grp * n_tiles;
    if (blocks <= 0) return hipSuccess;
    unsigned sg = 0;
    hipError_t e = spill_begin(A, blocks, stream, sg);
    if (e != hipSuccess) return e;
    if (exact) {
        hipLaunchKernelGGL(pe_backproject_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A, y, grp,
                           partial);
        if (sg) hipLaunchKernelGGL((pe_project_spill_kernel<true, true>), dim3(sg), dim3(64), 0, stream, A, y, grp,
                           partial);
    } else if (A.ell) {
        hipLaunchKernelGGL((pe_backproject_kernel<false, true>), dim3((unsigned)blocks), dim3(64), 0, stream, A, y,
                           grp, partial);
        if (sg)
            hipLaunchKernelGGL((pe_project_spill_kernel<false, true, true>), dim3(sg), dim3(64), 0, stream, A, y,
                               grp, partial);
    } else {
        hipLaunchKernelGGL(pe_backproject_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A, y, grp,
                           partial);
        if (sg) hipLaunchKernelGGL((pe_project_spill_kernel<false, true>), dim3(sg), dim3(64), 0, stream, A, y, grp,
                           partial);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = g.n_vox;  // x[v] = sum over groups of partial[k * V + v]
    hipLaunchKernelGGL(sum_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, partial,
                       n, n_grp, n, (int64_t)1, x);
    return hipGetLastError();
}

}  // namespace smg
