// Host build of the GPU lane arithmetic, for CPU tests against the oracle.
//
// Compiles the device headers of the product (csrc/smg_pe_body.hpp, ...) with
// g++ and evaluates PE elements one lane at a time: the reference culling
// predicate over every bar and hole (FP32 with the FP64 re-check, as on the
// GPU), then the sub-cell rays in FP32 with the FP64 recompute of flagged
// rays, or everything in FP64.  The GPU kernel adds the wave-level candidate
// narrowing, which is exact by construction and is checked on the GPU.
#include <cstdint>
#include <string>
#include <vector>

#include "smg_parse.hpp"
#include "smg_pe_body.hpp"

using namespace smg;

namespace {

struct VecCands {
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

bool make_ctx(const float* coll, const float* det, const float* img, const float* phys, Ctx& c) {
    std::string why;
    if (parse_geometry(coll, det, img, phys, c.hg, why) != 0) return false;
    c.G.g = c.hg.g;
    c.G.bars_f = c.hg.bars_f.data();
    c.G.bars_d = c.hg.bars_d.data();
    c.G.holes_f = c.hg.holes_f.data();
    c.G.holes_d = c.hg.holes_d.data();
    c.G.layers_f = c.hg.layers_f.data();
    c.G.layers_d = c.hg.layers_d.data();
    return true;
}

template <bool kExact>
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
                T val = pe_cell_ray<T>(G, d, p, ps, cc, C, near);
                if (!kExact && near) {
                    bool n2 = false;
                    val = (T)pe_cell_ray<double>(G, d, p, ps, cc, C, n2);
                    ++*fallbacks;
                }
                acc.add(val);
            }
    return (double)(float)acc.value();
}

}  // namespace

extern "C" {

// out[i] = A[d[i]][v[i]] of rotation rot; exact = 1 runs everything in FP64.
// Returns the number of FP64 fallback rays, or -1 on a geometry error.
long lane_pe_pairs(const float* coll, const float* det, const float* img, const float* phys,
                   int rot, int sx, int sy, int sz, int exact, const int* d, const int* v, int n,
                   double* out) {
    Ctx c;
    if (!make_ctx(coll, det, img, phys, c)) return -1;
    long fb = 0;
    for (int i = 0; i < n; ++i)
        out[i] = exact ? element<true>(c, d[i], v[i], rot, sx, sy, sz, &fb)
                       : element<false>(c, d[i], v[i], rot, sx, sy, sz, &fb);
    return fb;
}

}  // extern "C"
