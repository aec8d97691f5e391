// Host build of the GPU lane arithmetic, for CPU tests against the oracle.
//
// Compiles the device headers of the product (csrc/smg_pe_body.hpp, ...) with
// g++ and evaluates PE elements one lane at a time: the reference culling
// predicate over every bar and hole (FP32 with the FP64 re-check, as on the
// GPU), then the sub-cell rays in FP32 with the FP64 recompute of flagged
// rays, or everything in FP64.  The GPU kernel adds the wave-level candidate
// narrowing, which is exact by construction and is checked on the GPU.
#include <algorithm>
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "smg_parse.hpp"
#include "smg_pe_body.hpp"
#include "smg_bitmap_body.hpp"
#include "smg_scatter_body.hpp"

using namespace smg;

namespace {

struct VecCands {
    static constexpr bool kFastFaces = false;
    std::vector<int> bars, holes;
    int nbars() const { return (int)bars.size(); }
    int bar(int j) const { return bars[j]; }
    bool bar_on(int) const { return true; }
    int nholes() const { return (int)holes.size(); }
    int hole(int j) const { return holes[j]; }
    bool hole_on(int) const { return true; }
};

struct Ctx {
    HostGeom hg;
    PeGeom G;
};

// Hole shapes applied to every geometry the harness builds (empty: circles).
std::vector<float> g_hole_shapes;

bool make_ctx(const float* coll, const float* det, const float* img, const float* phys, Ctx& c) {
    std::string why;
    if (parse_geometry(coll, det, img, phys, c.hg, why) != 0) return false;
    if (!g_hole_shapes.empty() &&
        apply_hole_shapes(c.hg, g_hole_shapes.data(), (int)g_hole_shapes.size() / 3, why) != 0)
        return false;
    c.G.g = c.hg.g;
    c.G.bars_f = c.hg.bars_f.data();
    c.G.bars_d = c.hg.bars_d.data();
    c.G.holes_f = c.hg.holes_f.data();
    c.G.holes_d = c.hg.holes_d.data();
    c.G.layers_f = c.hg.layers_f.data();
    c.G.layers_d = c.hg.layers_d.data();
    c.G.bar_groups = c.hg.bar_groups.data();
    c.G.hole_groups = c.hg.hole_groups.data();
    return true;
}

template <bool kExact, bool kEll>
double element(const Ctx& c, int d, int v, int rot, int sx, int sy, int sz, long* fallbacks) {
    const PeGeom& G = c.G;
    const GeomDesc& g = G.g;
    const Bar<double>& bd = G.bars_d[d];
    if (!(bd.mut > (double)0.01f)) return 0.0;
    int nxy = g.nx * g.ny;
    int iz = v / nxy, rem = v % nxy, iy = rem / g.nx, ix = rem % g.nx;
    double phi = (double)rot * g.dphi;
    double p[3], ps[3];
    voxel_position(g, ix, iy, iz, std::cos(phi), std::sin(phi), p);
    to_frame(bd, p, ps);
    float pf[3] = {float(p[0]), float(p[1]), float(p[2])};
    SelfCull sc;
    self_cull(bd, p, sc);
    float udf[3] = {float(sc.ud[0]), float(sc.ud[1]), float(sc.ud[2])};
    float critf = float(sc.crit);
    VecCands C;
    for (int k = 0; k < g.n_det; ++k)
        if (k != d && cull_bar<kExact>(G, k, p, pf, sc, udf, critf)) C.bars.push_back(k);
    for (int h = 0; h < g.n_holes; ++h)
        if (cull_hole<kExact>(G, h, p, pf, sc, udf, critf)) C.holes.push_back(h);
    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    for (int cz = 0; cz < sz; ++cz)
        for (int a = 0; a < sx; ++a)
            for (int b = 0; b < sy; ++b) {
                CellCtx cc;
                make_cell_ctx(bd, a, b, cz, sx, sy, sz, cc);
                bool near = false;
                T val = pe_cell_ray<T, kEll>(G, d, p, ps, cc, C, near);
                if (!kExact && near) {
                    bool n2 = false;
                    val = (T)pe_cell_ray<double, kEll>(G, d, p, ps, cc, C, n2);
                    ++*fallbacks;
                }
                acc.add(val);
            }
    return (double)(float)acc.value();
}

// Candidate lists narrowed for one sub-cell: a wave-uniform list of bars and
// holes with per-lane masks, as the GPU kernel holds them in LDS.
struct MaskCands {
    static constexpr bool kFastFaces = false;
    std::vector<int> b, h;
    std::vector<uint64_t> bm, hm;
    int lane = 0;
    int nbars() const { return (int)b.size(); }
    int bar(int j) const { return b[j]; }
    bool bar_on(int j) const { return (bm[j] >> lane) & 1ull; }
    int nholes() const { return (int)h.size(); }
    int hole(int j) const { return h[j]; }
    bool hole_on(int j) const { return (hm[j] >> lane) & 1ull; }
};

// MaskCands with the face records and mu_total of the first 32 narrowed bars,
// as the GPU kernels keep them in LDS (fast_faces), so pe_cell_ray<float>
// takes fast_bar_chord for the bars that qualify.
struct FastMaskCands : MaskCands {
    static constexpr bool kFastFaces = true;
    std::vector<char> ok;
    std::vector<std::array<float, 7>> f;
    bool fast(int j) const { return j < 32 && j < (int)ok.size() && ok[j]; }
    const float* faces(int j) const { return f[j].data(); }
    bool has_mu(int j) const { return j < (int)f.size(); }
    float mu(int j) const { return f[j][6]; }
};

}  // namespace

extern "C" {

// Length of the segment a -> b inside the elliptic cylinder (seg_ell_geo, the
// grid scatterers' collimator material), for a hole of semi-axes erx, erz
// turned by theta.
double lane_seg_ell(const double* a, const double* b, double xc, double y1, double y2, double zc,
                    double erx, double erz, double theta) {
    Hole<double> o{};
    o.xc = xc;
    o.y1 = y1;
    o.y2 = y2;
    o.zc = zc;
    o.erx = erx;
    o.erz = erz;
    o.ec = std::cos(theta);
    o.es = std::sin(theta);
    o.r = std::max(erx, erz);
    return seg_ell_geo(a, b, o);
}

// Work items of the blocks 0 .. n-1 of a PE launch (xcd_item) into out.
void lane_xcd_items(int n, int* out) {
    for (int b = 0; b < n; ++b) out[b] = xcd_item(b, n);
}

// Checks the wave-level candidate selection (bar_tile_may_hit per tile on
// group boxes and members, then box_meets_frustum per sub-cell)
// and the tile-level plate-gate skip (plate_gate_clear) on n_waves seeded
// (bar, tile) pairs: every lane's FP64 sub-cell term with the narrowed lists
// must equal, bit for bit, the term with the lane's full culling set.
// stats[0] = mismatching (lane, sub-cell) terms, stats[1] = terms checked,
// stats[2] = mean narrowed bars per sub-cell, stats[3] = mean narrowed holes,
// stats[4] = mean bars in the full sets, stats[5] = mean tile candidates
// (bars) kept by a culling test of some lane, stats[6] = share of (sub-cell,
// layer) plate gates skipped; the narrowed terms are computed with those
// gates skipped; stats[7] = share of (wave, layer) gates cleared for the
// whole bar (plate_gate_clear_ball), the rest tested per sub-cell;
// stats[8], stats[9] = the largest tile candidate lists (bars, holes) over the
// waves, which the kernels hold in LDS up to 448 bars and 256 holes;
// stats[10] = the most narrowed bars of one sub-cell, stats[11] = the share of
// sub-cells with more than 8 (the adjoint keeps the chords of the first 8 in
// LDS and recomputes the others); the FP32 term with the narrowed lists and
// the face records of fast_bar_chord against the FP64 term: stats[12] = the
// largest relative difference over the terms it does not flag, stats[13] =
// the share of narrowed bars (the first 32) that qualify, stats[14] = the
// share of flagged terms.
int lane_narrow_check(const float* coll, const float* det, const float* img, const float* phys,
                      int rot, int sx, int sy, int sz, int tx, int ty, int tz, int n_waves,
                      unsigned seed, double* stats) {
    Ctx c;
    if (!make_ctx(coll, det, img, phys, c)) return -1;
    const PeGeom& G = c.G;
    const GeomDesc& g = G.g;
    const int tile[3] = {tx, ty, tz};
    const int nt[3] = {(g.nx + tx - 1) / tx, (g.ny + ty - 1) / ty, (g.nz + tz - 1) / tz};
    const double phi = (double)rot * g.dphi, cphi = std::cos(phi), sphi = std::sin(phi);
    uint64_t st = seed * 2654435761ull + 1;
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    double bad = 0, checked = 0, nsb = 0, nsh = 0, nfull = 0, cells = 0, ntile = 0, nwav = 0, gates = 0,
           bar_gates = 0, max_tb = 0, max_th = 0, max_nsb = 0, over8 = 0, max_fast = 0, n_fast = 0,
           n_fastable = 0, n_near = 0;
    for (int w = 0; w < n_waves; ++w) {
        int d = (int)(next() % g.n_det);
        int t = (int)(next() % ((uint64_t)nt[0] * nt[1] * nt[2]));
        int ix0 = (t % nt[0]) * tx, iy0 = ((t / nt[0]) % nt[1]) * ty, iz0 = (t / (nt[0] * nt[1])) * tz;
        const Bar<double>& bd = G.bars_d[d];
        if (!(bd.mut > (double)0.01f)) continue;
        nwav += 1;
        double tc[3], rho;
        tile_ball(g, ix0, iy0, iz0, tile, cphi, sphi, tc, rho);
        float tcf[3], rhof;
        tile_ball_f(tc, rho, tcf, rhof);
        // plate gates cleared for the whole bar (as the kernel's wave set-up)
        uint32_t bar_gate = 0;
        {
            const double qc[3] = {bd.cx, bd.cy, bd.cz};
            const double rb = 0.5 * bd.rdiag * (1.0 + 1e-12) + 1e-9;
            for (int m = 0; m < g.n_layers; ++m)
                if (plate_gate_clear_ball(G.layers_d[m], qc, rb, tc, rho)) bar_gate |= 1u << m;
            bar_gates += __builtin_popcount(bar_gate);
        }
        // full culling sets (every bar and hole), and the kernel's tile
        // candidates (bar_tile_may_hit first, then the culling test)
        const Bar<float>& bdf = G.bars_f[d];
        std::vector<char> tb(g.n_det, 0), th(g.n_holes, 0);
        auto group_hit = [&](const GroupBox& B) {
            return bar_tile_may_hit(bdf, tcf, rhof, B.c[0], B.c[1], B.c[2], 1.0f, 0.0f, B.h[0],
                                    B.h[1], B.h[2]);
        };
        for (int k = 0; k < g.n_det; ++k) {
            const Bar<float>& b = G.bars_f[k];
            tb[k] = k != d && group_hit(G.bar_groups[k / kGroup]) &&
                    bar_tile_may_hit(bdf, tcf, rhof, b.cx, b.cy, b.cz, b.cb, b.sb, 0.5f * b.w,
                                     0.5f * b.t, 0.5f * b.h);
        }
        for (int h = 0; h < g.n_holes; ++h) {
            const Hole<float>& o = G.holes_f[h];
            th[h] = group_hit(G.hole_groups[h / kGroup]) &&
                    bar_tile_may_hit(bdf, tcf, rhof, o.xc, o.ymid, o.zc, 1.0f, 0.0f, o.r,
                                     0.5f * tabs(o.y2 - o.y1), o.r);
        }
        std::vector<uint64_t> bmask(g.n_det, 0), hmask(g.n_holes, 0);
        double P[64][3], PS[64][3];
        bool valid[64];
        for (int l = 0; l < tx * ty * tz; ++l) {
            int ix = ix0 + l % tx, iy = iy0 + (l / tx) % ty, iz = iz0 + l / (tx * ty);
            valid[l] = ix < g.nx && iy < g.ny && iz < g.nz;
            if (!valid[l]) continue;
            voxel_position(g, ix, iy, iz, cphi, sphi, P[l]);
            to_frame(bd, P[l], PS[l]);
            SelfCull sc;
            self_cull(bd, P[l], sc);
            float pf[3] = {float(P[l][0]), float(P[l][1]), float(P[l][2])};
            float udf[3] = {float(sc.ud[0]), float(sc.ud[1]), float(sc.ud[2])};
            for (int k = 0; k < g.n_det; ++k)
                if (k != d && cull_bar<true>(G, k, P[l], pf, sc, udf, float(sc.crit))) {
                    bmask[k] |= 1ull << l;
                    nfull += 1;
                }
            for (int h = 0; h < g.n_holes; ++h)
                if (cull_hole<true>(G, h, P[l], pf, sc, udf, float(sc.crit))) hmask[h] |= 1ull << l;
        }
        double wtb = 0, wth = 0;
        for (int k = 0; k < g.n_det; ++k) wtb += tb[k] && bmask[k];
        for (int h = 0; h < g.n_holes; ++h) wth += th[h] && hmask[h];
        ntile += wtb;
        max_tb = std::max(max_tb, wtb);
        max_th = std::max(max_th, wth);
        for (int cz = 0; cz < sz; ++cz)
            for (int a = 0; a < sx; ++a)
                for (int b = 0; b < sy; ++b) {
                    CellCtx cc;
                    make_cell_ctx(bd, a, b, cz, sx, sy, sz, cc);
                    const float q[3] = {float(cc.q[0]), float(cc.q[1]), float(cc.q[2])};
                    const float rq[3] = {0.0f, 0.0f, 0.0f};
                    MaskCands full, nar;
                    for (int k = 0; k < g.n_det; ++k) {
                        if (!bmask[k]) continue;
                        full.b.push_back(k);
                        full.bm.push_back(bmask[k]);
                        const Bar<float>& bf = G.bars_f[k];
                        if (tb[k] && box_meets_frustum(q, rq, tcf, rhof, bf.cx, bf.cy, bf.cz, bf.cb, bf.sb,
                                                       0.5f * bf.w, 0.5f * bf.t, 0.5f * bf.h)) {
                            nar.b.push_back(k);
                            nar.bm.push_back(bmask[k]);
                        }
                    }
                    for (int h = 0; h < g.n_holes; ++h) {
                        if (!hmask[h]) continue;
                        full.h.push_back(h);
                        full.hm.push_back(hmask[h]);
                        const Hole<float>& o = G.holes_f[h];
                        if (th[h] && box_meets_frustum(q, rq, tcf, rhof, o.xc, o.ymid, o.zc, 1.0f, 0.0f,
                                                       o.r, 0.5f * tabs(o.y2 - o.y1), o.r)) {
                            nar.h.push_back(h);
                            nar.hm.push_back(hmask[h]);
                        }
                    }
                    nsb += nar.b.size();
                    nsh += nar.h.size();
                    max_nsb = std::max(max_nsb, (double)nar.b.size());
                    over8 += nar.b.size() > 8;
                    cells += 1;
                    // the same sub-cell with the tile's plate gates skipped
                    CellCtx cg = cc;
                    cg.gate_clear = bar_gate;
                    for (int m = 0; m < g.n_layers; ++m)
                        if (!((bar_gate >> m) & 1u) && plate_gate_clear(G.layers_d[m], cc.q, tc, rho))
                            cg.gate_clear |= 1u << m;
                    gates += __builtin_popcount(cg.gate_clear);
                    FastMaskCands fm;
                    static_cast<MaskCands&>(fm) = nar;
                    for (int j = 0; j < (int)nar.b.size() && j < 32; ++j) {
                        std::array<float, 7> fr;
                        fm.ok.push_back(fast_faces(G, nar.b[j], cc.q, tc, rho, fr.data()) ? 1 : 0);
                        fr[6] = G.bars_f[nar.b[j]].mut;
                        fm.f.push_back(fr);
                        n_fast += fm.ok.back();
                        n_fastable += 1;
                    }
                    for (int l = 0; l < tx * ty * tz; ++l) {
                        if (!valid[l]) continue;
                        full.lane = nar.lane = fm.lane = l;
                        bool n1 = false, n2 = false, n3 = false;
                        double v1 = pe_cell_ray<double, true>(G, d, P[l], PS[l], cc, full, n1);
                        double v2 = pe_cell_ray<double, true>(G, d, P[l], PS[l], cg, nar, n2);
                        checked += 1;
                        if (!(v1 == v2)) bad += 1;
                        double v3 = (double)pe_cell_ray<float, false>(G, d, P[l], PS[l], cg, fm, n3);
                        if (n3) {
                            n_near += 1;
                        } else if (v1 != 0.0 || v3 != 0.0) {
                            max_fast = std::max(max_fast, std::fabs(v3 - v1) / std::fabs(v1));
                        }
                    }
                }
    }
    stats[0] = bad;
    stats[1] = checked;
    stats[2] = cells > 0 ? nsb / cells : 0;
    stats[3] = cells > 0 ? nsh / cells : 0;
    stats[4] = checked > 0 ? nfull / (checked / (sx * sy * sz)) : 0;
    stats[5] = nwav > 0 ? ntile / nwav : 0;
    stats[6] = cells > 0 ? gates / (cells * (g.n_layers > 0 ? g.n_layers : 1)) : 0;
    stats[7] = nwav > 0 ? bar_gates / (nwav * (g.n_layers > 0 ? g.n_layers : 1)) : 0;
    stats[8] = max_tb;
    stats[9] = max_th;
    stats[10] = max_nsb;
    stats[11] = cells > 0 ? over8 / cells : 0;
    stats[12] = max_fast;
    stats[13] = n_fastable > 0 ? n_fast / n_fastable : 0;
    stats[14] = checked > 0 ? n_near / checked : 0;
    return 0;
}

// out[i] = A[d[i]][v[i]] of rotation rot; exact = 1 runs everything in FP64.
// Returns the number of FP64 fallback rays, or -1 on a geometry error.
long lane_pe_pairs(const float* coll, const float* det, const float* img, const float* phys,
                   int rot, int sx, int sy, int sz, int exact, const int* d, const int* v, int n,
                   double* out) {
    Ctx c;
    if (!make_ctx(coll, det, img, phys, c)) return -1;
    // as the library's matrix kernels do: the elliptic chord is compiled in
    // while a hole is elliptical
    bool ell = false;
    for (const Hole<double>& o : c.hg.holes_d) ell |= o.erx > 0.0;
    long fb = 0;
    for (int i = 0; i < n; ++i)
        out[i] = exact ? element<true, true>(c, d[i], v[i], rot, sx, sy, sz, &fb)
                 : ell ? element<false, true>(c, d[i], v[i], rot, sx, sy, sz, &fb)
                       : element<false, false>(c, d[i], v[i], rot, sx, sy, sz, &fb);
    return fb;
}

}  // extern "C"

namespace {

bool bit_set(const uint32_t* words, int64_t b) { return (words[b >> 5] >> (b & 31)) & 1u; }

// Pair terms of one (d, scatter point, sub-cell), as the pair-terms kernel forms them.
PairTerm pair_term(const Ctx& c, const uint32_t* bitmap, int kind, int d, int s, const double qs[3],
                   const CellCtx& cc) {
    const PeGeom& G = c.G;
    const int N = G.g.n_det;
    const Bar<double>& bd = G.bars_d[d];
    PairTerm out{0, 0, 0, 0};
    bool near = false;
    int64_t base;
    if (kind == 0) {
        int64_t i = s < d ? s : d, j = s < d ? d : s;
        base = (i * N - i * (i + 1) / 2 + j) * N;
    } else {
        base = ((int64_t)s * N + d) * N;
    }
    for (int k = 0; k < N; ++k) {
        if (k == d || (kind == 0 && k == s) || !bit_set(bitmap, base + k)) continue;
        const Bar<double>& bk = G.bars_d[k];
        double pa[3], pb[3];
        to_frame(bk, qs, pa);
        to_frame(bk, cc.q, pb);
        double lo[3] = {-0.5 * bk.w, -0.5 * bk.t, -0.5 * bk.h};
        double hi[3] = {0.5 * bk.w, 0.5 * bk.t, 0.5 * bk.h};
        SegDir<double> sd;
        segdir_init<double>(sd, pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2], 0.0, near);
        double l = seg_box<double>(sd, pa, lo, hi, near);
        out.s_pe += l * bk.mupe;
        out.s_c += l * bk.muc;
    }
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
    double dx = qs[0] - cc.q[0], dy = qs[1] - cc.q[1], dz = qs[2] - cc.q[2];
    out.len = std::sqrt(dx * dx + dy * dy + dz * dz);
    double De[3];
    for (int k = 0; k < 3; ++k) De[k] = (cc.qs[k] + 1000.0 * (cc.qs[k] - qss[k]) / out.len) - qss[k];
    SegDir<double> se;
    segdir_init<double>(se, De[0], De[1], De[2], 0.0, near);
    out.la = seg_box<double>(se, qss, cc.lo, cc.hi, near);
    return out;
}

double kn_norm(double e0) {
    const double step = (double)0.01f, hi = (double)(float)3.14159265358979323846;
    int n = (int)(hi / step);
    double s = 0.0, alpha = e0 / 511.0;
    for (int i = 0; i <= n; ++i) {
        double th = i * step;
        s += kn_of_cos(std::cos(th), alpha) * std::sin(th) * step;
    }
    return s;
}

template <class T>
double scatter_element(const Ctx& c, const float* pe, const uint32_t* bitmap, int kind, int variant,
                       int d, int v, int rot, int sx, int sy, int sz, long* fallbacks) {
    const PeGeom& G = c.G;
    const GeomDesc& g = G.g;
    const HostGeom& hg = c.hg;
    const Bar<double>& bd = G.bars_d[d];
    const float* ph = hg.phys.data();
    if ((int)std::floor((double)ph[0] + 0.5) != 1 || !(bd.mut > (double)0.01f)) return 0.0;
    int nxy = g.nx * g.ny;
    int iz = v / nxy, rem = v % nxy;
    double phi = (double)rot * g.dphi;
    double p[3], ps[3];
    voxel_position(g, rem % g.nx, rem / g.nx, iz, std::cos(phi), std::sin(phi), p);
    to_frame(bd, p, ps);
    ScatterPhys P;
    P.e0 = ph[7];
    if ((int)std::floor((double)ph[4] + 0.5) > 0) {
        P.lo = ph[5];
        P.hi = ph[6];
    } else {
        P.lo = (1 - bd.res / 2.0) * P.e0;
        P.hi = (1 + bd.res / 2.0) * P.e0;
    }
    P.norm = kn_norm(P.e0);
    P.res = bd.res;
    P.mupe = bd.mupe;
    P.muc = bd.muc;
    P.rcell = std::sqrt(bd.w * bd.w / sx / sx + bd.h * bd.h / sz / sz + bd.t * bd.t / sy / sy) / 2.0;
    // the row cut at margin 0 (row_cut_kernel), then the path's margin
    P.c_rej = window_cut_margin(window_reject_cos(P, 0.0), std::is_same<T, double>::value ? 1e-9 : 1e-5);
    // the KN prefix table the GPU's FP32 path reads (built with the same code)
    std::vector<double> knt((size_t)(kKnPhase + 1) * kKnSteps);
    for (int j = 0; j <= kKnPhase; ++j) kn_prefix_row(P.e0 / 511.0, j, knt.data() + (size_t)j * kKnSteps);
    P.knt = knt.data();
    int n_pts = kind == 0 ? g.n_det : hg.n_holes_l0;
    KahanSum<T> acc;
    for (int cell = 0; cell < sx * sy * sz; ++cell) {
        int a = cell % sx, b = (cell / sx) % sy, cz = cell / (sx * sy);
        CellCtx cc;
        make_cell_ctx(bd, a, b, cz, sx, sy, sz, cc);
        ScatterCell sc;
        for (int k = 0; k < 3; ++k) {
            sc.q[k] = cc.q[k];
            sc.lo[k] = cc.lo[k];
            sc.hi[k] = cc.hi[k];
        }
        for (int s = 0; s < n_pts; ++s) {
            if (kind == 0 && s == d) continue;
            double dyh = 0.0;
            if (kind == 0) {
                const Bar<double>& bs = G.bars_d[s];
                sc.qs[0] = bs.cx;
                sc.qs[1] = bs.cy;
                sc.qs[2] = bs.cz;
            } else {
                const float* hp = hg.coll.data() + 100 + 9 * s;
                sc.qs[0] = hp[0];
                sc.qs[1] = ((double)hp[1] + (double)hp[2]) / 2.0 + g.fov2coll;
                sc.qs[2] = hp[3];
                dyh = (double)hp[2] - (double)hp[1];
            }
            to_frame(bd, sc.qs, sc.qss);
            sc.pt = pair_term(c, bitmap, kind, d, s, sc.qs, cc);
            T wgt;
            if (kind == 0) {
                wgt = T(pe[(int64_t)s * g.n_vox + v]) * T(bd.muc / bd.mupe);
            } else {
                double dx = p[0] - sc.qs[0], dy = p[1] - sc.qs[1], dz = p[2] - sc.qs[2];
                double lv = std::sqrt(dx * dx + dy * dy + dz * dz);
                double om;
                if (variant == 0) {
                    double r = (double)0.1f;
                    om = (r * r * 3.14159265358979323846 + dyh * r) /
                         (4.0 * 3.14159265358979323846 * lv * lv) * std::fabs(sc.qs[1] - p[1]) / lv;
                } else {
                    double r = (double)0.2f;
                    om = (r * r + dyh * r) / (4.0 * lv * lv);
                }
                wgt = T(om * ((double)hg.coll[17] / (double)hg.coll[15]));
            }
            if (wgt == T(0)) continue;
            bool near = false;
            T val = scatter_term<T>(P, p, ps, sc, wgt, near);
            if (std::is_same<T, float>::value && near) {
                bool n2 = false;
                val = T(scatter_term<double>(P, p, ps, sc, double(wgt), n2));
                ++*fallbacks;
            }
            acc.add(val);
        }
    }
    return (double)(float)acc.value();
}

}  // namespace

extern "C" {

// kind 0: crystal scatter (pe = N x V block), 1: collimator scatter (variant 0/1).
long lane_scatter_pairs(const float* coll, const float* det, const float* img, const float* phys,
                        const float* pe, const uint32_t* bitmap, int kind, int variant, int rot,
                        int sx, int sy, int sz, int exact, const int* d, const int* v, int n,
                        double* out) {
    Ctx c;
    if (!make_ctx(coll, det, img, phys, c)) return -1;
    long fb = 0;
    for (int i = 0; i < n; ++i)
        out[i] = exact ? scatter_element<double>(c, pe, bitmap, kind, variant, d[i], v[i], rot, sx,
                                                 sy, sz, &fb)
                       : scatter_element<float>(c, pe, bitmap, kind, variant, d[i], v[i], rot, sx,
                                                sy, sz, &fb);
    return fb;
}


// KN rectangle sums over [lo[i], hi[i]] from the prefix table of e0 (the FP32
// scatter path's kn_sum_table); near[i] = 1 where the term would go to FP64.
void lane_kn_sum_table(double e0, const float* lo, const float* hi, int n, double* out, int* near) {
    std::vector<double> knt((size_t)(kKnPhase + 1) * kKnSteps);
    for (int j = 0; j <= kKnPhase; ++j) kn_prefix_row(e0 / 511.0, j, knt.data() + (size_t)j * kKnSteps);
    for (int i = 0; i < n; ++i) {
        bool nr = false;
        out[i] = kn_sum_table(knt.data(), lo[i], hi[i], nr);
        near[i] = nr;
    }
}

// mu(E')/mu(E0) of the four tables at E' = ep[i]: out32 from the FP32 form
// the FP32 scatter term uses (piecewise power law in ln(E0/E')), out64 from the
// FP64 table search, 4 values per point (crystal pe, crystal Compton,
// collimator pe, collimator Compton).  Returns the FP32 form's segment count
// (0: not available).
int lane_mu_ratio_f(double e0, int n_nodes, const double* energy, const double* cpe, const double* cc,
                    const double* wpe, const double* wc, const double* ep, int n, double* out32,
                    double* out64) {
    MuTables t{};
    t.n = n_nodes;
    const double* v[4] = {cpe, cc, wpe, wc};
    for (int i = 0; i < n_nodes; ++i) {
        t.le[i] = std::log(energy[i]);
        for (int k = 0; k < 4; ++k) t.lmu[k][i] = std::log(v[k][i]);
    }
    for (int k = 0; k < 4; ++k) t.lmu0[k] = mu_table_log(t, k, std::log(e0));
    MuTablesF f{};
    make_mu_tables_f(t, e0, f);
    const double alpha = e0 / 511.0;
    for (int i = 0; i < n; ++i) {
        const float L = logf(float(1.0 + alpha * (e0 / ep[i] - 1.0) / alpha));
        const int j = f.n > 0 ? mu_seg_f(f, L) : 0;
        for (int k = 0; k < 4; ++k) {
            out64[4 * i + k] = std::exp(mu_table_log(t, k, std::log(ep[i])) - t.lmu0[k]);
            out32[4 * i + k] = f.n > 0 ? (double)mu_ratio_f(f, k, j, L) : 0.0;
        }
    }
    return f.n;
}

// The FP32 scatter term's lower-window threshold on cos(theta).
double lane_window_reject_cos(double e0, double lo, double res, double margin) {
    ScatterPhys P;
    P.e0 = e0;
    P.lo = lo;
    P.hi = 2.0 * e0;
    P.res = res;
    return window_reject_cos(P, margin);
}

// The scatter kernel's threshold from the row cut at margin 0.
double lane_window_cut_margin(double cut, double margin) { return window_cut_margin(cut, margin); }

// Elliptical holes (x_rho, z_rho, theta per hole) for the following calls;
// n = 0 restores circles.
void lane_set_hole_shapes(const float* shapes, int n) {
    g_hole_shapes.assign(shapes, shapes + 3 * (size_t)n);
}

// Crystal (collimator = 0) or collimator bitmap words by the GPU kernel's
// walk: runs of 64 words, lane l of a run walking bits 64 s + l (s = 0..31)
// with its row terms kept across steps.  Returns the number of words, or -1.
long lane_bitmap(const float* coll, const float* det, const float* img, const float* phys,
                 int collimator, uint32_t* words, long n_words) {
    HostGeom hg;
    std::string why;
    if (parse_geometry(coll, det, img, phys, hg, why) != 0) return -1;
    std::vector<Sphere<float>> sf, hf;
    std::vector<Sphere<double>> sd, hd;
    std::vector<ScatterPt<double>> pb, ph;
    make_scatter_records(hg, sf, sd, hf, hd, pb, ph);
    const bool cl = collimator != 0;
    const int64_t n = hg.g.n_det;
    const int64_t total = bm_total_bits(cl, n, hg.n_holes_l0);
    const long need = (long)((total + 31) / 32);
    if (n_words < need) return -1;
    for (long w = 0; w < need; ++w) words[w] = 0;
    const int64_t n_runs = (need + 63) / 64;
    for (int64_t run = 0; run < n_runs; ++run) {
        for (int lane = 0; lane < 64; ++lane) {
            int64_t b = run * 2048 + lane;
            BitRow r{};
            int k = 0;
            if (b < total) {
                bm_row_of_bit(cl, b, n, r, k);
                bm_row_setup(cl, sd.data(), hd.data(), r);
            }
            for (int s = 0; s < 32; ++s) {
                if (b < total && bm_bit(cl, r, sd.data(), hd.data(), k)) {
                    const int64_t w = run * 64 + 2 * s + lane / 32;
                    words[w] |= 1u << (lane % 32);
                }
                b += 64;
                k += 64;
                if (b < total && k >= n) {
                    do {
                        k -= (int)n;
                        bm_next_row(cl, n, r);
                    } while (k >= n);
                    bm_row_setup(cl, sd.data(), hd.data(), r);
                }
            }
        }
    }
    return need;
}

}  // extern "C"
