// Photoelectric system-matrix kernel for gfx950 (CDNA4, wave64).
//
// Work decomposition: one wave per (detector bar d, tile of 64 voxels).  The
// reference runs one thread per (d, v) and makes every thread scan all bars
// and holes for culling (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1097-1186)
// before its sub-cell loop (:1193-1362).  Here the 64 lanes of a wave share
// detector d, so every sub-cell point is wave-uniform and:
//   1. the wave scans the bars and holes once, 64 at a time, with a
//      conservative test against the tile's bounding sphere; each survivor is
//      then tested with the reference culling predicate per lane and stored
//      in LDS with the 64-bit ballot of the lanes that keep it;
//   2. for each sub-cell, the candidates are narrowed to those whose bounding
//      sphere can meet the cone from the sub-cell centre to the tile (a bar
//      outside that cone has chord 0 for every lane, so dropping it is exact);
//   3. every lane traces its own ray against the narrowed lists (wave-uniform
//      loops, per-lane masks) and accumulates the sub-cell terms with a
//      two-sum compensated sum;
//   4. each lane writes its element once: no atomics, deterministic results.
#include <hip/hip_runtime.h>

#include "smg_kernels.hpp"
#include "smg_pe_body.hpp"

namespace smg {

namespace {

// Candidate lists of one wave, in LDS.
struct WaveLists {
    uint64_t bar_mask[kCapBars];
    uint64_t hole_mask[kCapHoles];
    uint16_t bar_idx[kCapBars];
    uint16_t hole_idx[kCapHoles];
    uint16_t bar_sel[kCapBars];
    uint16_t hole_sel[kCapHoles];
};

// View of the narrowed lists for one lane.
struct LdsCands {
    const WaveLists* L;
    int nb, nh, lane;
    __device__ int nbars() const { return nb; }
    __device__ int bar(int j) const { return L->bar_idx[L->bar_sel[j]]; }
    __device__ bool bar_on(int j) const { return (L->bar_mask[L->bar_sel[j]] >> lane) & 1ull; }
    __device__ int nholes() const { return nh; }
    __device__ int hole(int j) const { return L->hole_idx[L->hole_sel[j]]; }
    __device__ bool hole_on(int j) const { return (L->hole_mask[L->hole_sel[j]] >> lane) & 1ull; }
};

__device__ __forceinline__ int lanes_below(uint64_t m, int lane) {
    return __popcll(m & ((1ull << lane) - 1ull));
}

// Conservative superset of the reference culling test for any voxel in the
// ball (tc, rho): the angle between the directions to a fixed point seen from
// two points of the ball differs by at most asin(rho/D) <= rho/sqrt(D^2-rho^2),
// and the thresholds grow at most to r/(D - rho).
__device__ __forceinline__ bool tile_may_keep(const double tc[3], double rho, const double ud[3],
                                              double Dd, double r_self, double cx, double cy,
                                              double cz, double r_other) {
    double dx = cx - tc[0], dy = cy - tc[1], dz = cz - tc[2];
    double D = sqrt(dx * dx + dy * dy + dz * dz);
    if (D <= 1.01 * rho || Dd <= 1.01 * rho) return true;
    double ux = dx / D - ud[0], uy = dy / D - ud[1], uz = dz / D - ud[2];
    double s = sqrt(ux * ux + uy * uy + uz * uz);
    double e = rho / sqrt(D * D - rho * rho) + rho / sqrt(Dd * Dd - rho * rho);
    double lim = r_other / (D - rho) + r_self / (Dd - rho);
    return s - e <= lim + 1e-4;
}

// Can any ray from a point of the ball (tc, rho) to the apex q touch the
// sphere (c, R)?  The rays lie in the cone of half-angle asin(rho/Lt) around
// q -> tc, cut at Lt + rho.
struct Cone {
    float q[3], u[3];
    float Lt, rho, tana, seca;
    bool all;
};

__device__ __forceinline__ void make_cone(Cone& k, const double q[3], const double tc[3], double rho) {
    double ux = tc[0] - q[0], uy = tc[1] - q[1], uz = tc[2] - q[2];
    double Lt = sqrt(ux * ux + uy * uy + uz * uz);
    k.all = Lt <= 1.01 * rho;
    double sa = k.all ? 0.0 : rho / Lt;
    double ca = sqrt(1.0 - sa * sa);
    for (int i = 0; i < 3; ++i) k.q[i] = float(q[i]);
    k.u[0] = float(ux / Lt);
    k.u[1] = float(uy / Lt);
    k.u[2] = float(uz / Lt);
    k.Lt = float(Lt);
    k.rho = float(rho);
    k.tana = float(sa / ca);
    k.seca = float(1.0 / ca);
}

__device__ __forceinline__ bool cone_hits(const Cone& k, float cx, float cy, float cz, float R) {
    if (k.all) return true;
    float wx = cx - k.q[0], wy = cy - k.q[1], wz = cz - k.q[2];
    float proj = wx * k.u[0] + wy * k.u[1] + wz * k.u[2];
    float w2 = wx * wx + wy * wy + wz * wz;
    float perp = sqrtf(fmaxf(w2 - proj * proj, 0.0f));
    float m = 2e-3f + 1e-5f * sqrtf(w2);
    return proj >= -R - m && proj <= k.Lt + k.rho + R + m &&
           perp <= R * k.seca + fmaxf(proj, 0.0f) * k.tana + m;
}

template <bool kExact>
__global__ __launch_bounds__(64) void pe_forward_kernel(PeLaunch A) {
    __shared__ WaveLists L;
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int d = (int)(blockIdx.x / n_tiles);
    const int tile = (int)(blockIdx.x % n_tiles);
    const int t0 = tile % A.tiles[0];
    const int t1 = (tile / A.tiles[0]) % A.tiles[1];
    const int t2 = tile / (A.tiles[0] * A.tiles[1]);
    const int ix0 = t0 * A.tile[0], iy0 = t1 * A.tile[1], iz0 = t2 * A.tile[2];
    const int ix = ix0 + lane % A.tile[0];
    const int iy = iy0 + (lane / A.tile[0]) % A.tile[1];
    const int iz = iz0 + lane / (A.tile[0] * A.tile[1]);
    const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
    const int64_t v = (int64_t)ix + (int64_t)iy * g.nx + (int64_t)iz * g.nx * g.ny;
    float* out = A.dst + (int64_t)d * g.n_vox;

    const Bar<double>& bd = G.bars_d[d];
    if (!(bd.mut > (double)0.01f)) {  // PESysMatGen.cu:1116, :1194
        if (valid) out[v] = 0.0f;
        return;
    }

    double p[3];
    voxel_position(g, valid ? ix : ix0, valid ? iy : iy0, valid ? iz : iz0, A.cphi, A.sphi, p);
    double ps[3];
    to_frame(bd, p, ps);
    const float pf[3] = {float(p[0]), float(p[1]), float(p[2])};
    SelfCull sc;
    self_cull(bd, p, sc);
    const float udf[3] = {float(sc.ud[0]), float(sc.ud[1]), float(sc.ud[2])};
    const float critf = float(sc.crit);

    // tile bounding sphere (the rotation about z preserves distances)
    const int ex = min(A.tile[0], g.nx - ix0), ey = min(A.tile[1], g.ny - iy0),
              ez = min(A.tile[2], g.nz - iz0);
    double tc[3];
    {
        double x = (ix0 + 0.5 * (ex - 1) - g.nx / 2.0 + 0.5) * g.wx + g.sx;
        double y = (iy0 + 0.5 * (ey - 1) - g.ny / 2.0 + 0.5) * g.wy + g.sy;
        double z = (iz0 + 0.5 * (ez - 1) - g.nz / 2.0 + 0.5) * g.wz + g.sz;
        tc[0] = x * A.cphi - y * A.sphi;
        tc[1] = x * A.sphi + y * A.cphi;
        tc[2] = z;
    }
    const double rho = 0.5 * sqrt(((ex - 1) * g.wx) * ((ex - 1) * g.wx) +
                                  ((ey - 1) * g.wy) * ((ey - 1) * g.wy) +
                                  ((ez - 1) * g.wz) * ((ez - 1) * g.wz)) + 1e-6;
    double ud_t[3], Dd_t;
    {
        double dx = bd.cx - tc[0], dy = bd.cy - tc[1], dz = bd.cz - tc[2];
        Dd_t = sqrt(dx * dx + dy * dy + dz * dz);
        ud_t[0] = dx / Dd_t;
        ud_t[1] = dy / Dd_t;
        ud_t[2] = dz / Dd_t;
    }
    const double rself = 0.5 * bd.rdiag;

    // ---- 1. candidate bars and holes with per-lane reference masks ----------------
    int nb = 0, nh = 0;
    for (int base = 0; base < g.n_det; base += 64) {
        int k = base + lane;
        bool pass = false;
        if (k < g.n_det && k != d) {
            const Bar<double>& b = G.bars_d[k];
            pass = tile_may_keep(tc, rho, ud_t, Dd_t, rself, b.cx, b.cy, b.cz, 0.5 * b.rdiag);
        }
        uint64_t bal = __ballot(pass);
        while (bal) {
            int bit = __ffsll((unsigned long long)bal) - 1;
            bal &= bal - 1;
            int kk = base + bit;
            bool keep = valid && cull_bar<kExact>(G, kk, p, pf, sc, udf, critf);
            uint64_t m = __ballot(keep);
            if (m) {
                if (nb < kCapBars && lane == 0) {
                    L.bar_idx[nb] = (uint16_t)kk;
                    L.bar_mask[nb] = m;
                }
                ++nb;
            }
        }
    }
    for (int base = 0; base < g.n_holes; base += 64) {
        int h = base + lane;
        bool pass = false;
        if (h < g.n_holes) {
            const Hole<double>& o = G.holes_d[h];
            pass = tile_may_keep(tc, rho, ud_t, Dd_t, rself, o.xc, o.ymid, o.zc, o.crit);
        }
        uint64_t bal = __ballot(pass);
        while (bal) {
            int bit = __ffsll((unsigned long long)bal) - 1;
            bal &= bal - 1;
            int hh = base + bit;
            bool keep = valid && cull_hole<kExact>(G, hh, p, pf, sc, udf, critf);
            uint64_t m = __ballot(keep);
            if (m) {
                if (nh < kCapHoles && lane == 0) {
                    L.hole_idx[nh] = (uint16_t)hh;
                    L.hole_mask[nh] = m;
                }
                ++nh;
            }
        }
    }
    if (nb > kCapBars || nh > kCapHoles) {
        // too many candidates for the LDS lists: report, write NaN
        if (lane == 0) atomicAdd(&A.counters[1], 1ull);
        if (valid) out[v] = __builtin_nanf("");
        return;
    }
    __syncthreads();

    // ---- 2./3. sub-cell loop --------------------------------------------------------
    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    unsigned fallbacks = 0;
    for (int c = 0; c < A.sub[2]; ++c)
        for (int a = 0; a < A.sub[0]; ++a)
            for (int b = 0; b < A.sub[1]; ++b) {
                CellCtx cc;
                make_cell_ctx(bd, a, b, c, A.sub[0], A.sub[1], A.sub[2], cc);
                Cone cone;
                make_cone(cone, cc.q, tc, rho);
                int nsb = 0, nsh = 0;
                for (int base = 0; base < nb; base += 64) {
                    int j = base + lane;
                    bool pass = false;
                    if (j < nb) {
                        const Bar<float>& bf = G.bars_f[L.bar_idx[j]];
                        pass = cone_hits(cone, bf.cx, bf.cy, bf.cz, 0.5f * bf.rdiag);
                    }
                    uint64_t bal = __ballot(pass);
                    if (pass) L.bar_sel[nsb + lanes_below(bal, lane)] = (uint16_t)j;
                    nsb += __popcll(bal);
                }
                for (int base = 0; base < nh; base += 64) {
                    int j = base + lane;
                    bool pass = false;
                    if (j < nh) {
                        const Hole<float>& o = G.holes_f[L.hole_idx[j]];
                        pass = cone_hits(cone, o.xc, o.ymid, o.zc, o.crit);
                    }
                    uint64_t bal = __ballot(pass);
                    if (pass) L.hole_sel[nsh + lanes_below(bal, lane)] = (uint16_t)j;
                    nsh += __popcll(bal);
                }
                __syncthreads();
                if (valid) {
                    LdsCands C{&L, nsb, nsh, lane};
                    bool near = false;
                    T val = pe_cell_ray<T>(G, d, p, ps, cc, C, near);
                    if constexpr (!kExact) {
                        if (near) {
                            bool n2 = false;
                            val = (T)pe_cell_ray<double>(G, d, p, ps, cc, C, n2);
                            ++fallbacks;
                        }
                    }
                    acc.add(val);
                }
                __syncthreads();
            }
    if (valid) out[v] = float(acc.value());
    if constexpr (!kExact) {
        // per-wave fallback count
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
}

}  // namespace

hipError_t launch_pe_forward(const PeLaunch& A, bool exact, hipStream_t stream) {
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int64_t blocks = (int64_t)n_tiles * A.geom.g.n_det;
    if (blocks <= 0) return hipSuccess;
    if (exact)
        hipLaunchKernelGGL(pe_forward_kernel<true>, dim3((unsigned)blocks), dim3(64), 0, stream, A);
    else
        hipLaunchKernelGGL(pe_forward_kernel<false>, dim3((unsigned)blocks), dim3(64), 0, stream, A);
    return hipGetLastError();
}

}  // namespace smg
