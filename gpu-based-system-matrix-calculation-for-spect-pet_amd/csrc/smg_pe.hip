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
    double tc[3], rho, ud_t[3], Dd_t, rself;
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
    // tile bounding sphere (the rotation about z preserves distances)
    const int ex = min(A.tile[0], g.nx - ix0), ey = min(A.tile[1], g.ny - iy0),
              ez = min(A.tile[2], g.nz - iz0);
    double x = (ix0 + 0.5 * (ex - 1) - g.nx / 2.0 + 0.5) * g.wx + g.sx;
    double y = (iy0 + 0.5 * (ey - 1) - g.ny / 2.0 + 0.5) * g.wy + g.sy;
    double z = (iz0 + 0.5 * (ez - 1) - g.nz / 2.0 + 0.5) * g.wz + g.sz;
    W.tc[0] = x * A.cphi - y * A.sphi;
    W.tc[1] = x * A.sphi + y * A.cphi;
    W.tc[2] = z;
    W.rho = 0.5 * sqrt(((ex - 1) * g.wx) * ((ex - 1) * g.wx) + ((ey - 1) * g.wy) * ((ey - 1) * g.wy) +
                       ((ez - 1) * g.wz) * ((ez - 1) * g.wz)) + 1e-6;
    double dx = bd.cx - W.tc[0], dy = bd.cy - W.tc[1], dz = bd.cz - W.tc[2];
    W.Dd_t = sqrt(dx * dx + dy * dy + dz * dz);
    W.ud_t[0] = dx / W.Dd_t;
    W.ud_t[1] = dy / W.Dd_t;
    W.ud_t[2] = dz / W.Dd_t;
    W.rself = 0.5 * bd.rdiag;
}

// Step 1: candidate bars and holes of the tile, each with the ballot of the
// lanes whose reference culling test keeps it.  Returns false on overflow.
template <bool kExact>
__device__ bool build_candidates(const PeLaunch& A, int d, int lane, const WaveSetup& W,
                                 WaveLists& L, int& nb, int& nh) {
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    nb = 0;
    nh = 0;
    for (int base = 0; base < g.n_det; base += 64) {
        int k = base + lane;
        bool pass = false;
        if (k < g.n_det && k != d) {
            const Bar<double>& b = G.bars_d[k];
            pass = tile_may_keep(W.tc, W.rho, W.ud_t, W.Dd_t, W.rself, b.cx, b.cy, b.cz, 0.5 * b.rdiag);
        }
        uint64_t bal = __ballot(pass);
        while (bal) {
            int bit = __ffsll((unsigned long long)bal) - 1;
            bal &= bal - 1;
            int kk = base + bit;
            bool keep = W.valid && cull_bar<kExact>(G, kk, W.p, W.pf, W.sc, W.udf, W.critf);
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
            pass = tile_may_keep(W.tc, W.rho, W.ud_t, W.Dd_t, W.rself, o.xc, o.ymid, o.zc, o.crit);
        }
        uint64_t bal = __ballot(pass);
        while (bal) {
            int bit = __ffsll((unsigned long long)bal) - 1;
            bal &= bal - 1;
            int hh = base + bit;
            bool keep = W.valid && cull_hole<kExact>(G, hh, W.p, W.pf, W.sc, W.udf, W.critf);
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
    return nb <= kCapBars && nh <= kCapHoles;
}

// Step 2: candidates whose bounding sphere can meet the cone from the
// sub-cell point to the tile; writes their list indices to bar_sel/hole_sel.
__device__ void narrow_candidates(const PeLaunch& A, const CellCtx& cc, const WaveSetup& W,
                                  int lane, WaveLists& L, int nb, int nh, int& nsb, int& nsh) {
    const PeGeom& G = A.geom;
    Cone cone;
    make_cone(cone, cc.q, W.tc, W.rho);
    nsb = 0;
    nsh = 0;
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
    float* out = A.dst + (int64_t)d * g.n_vox;
    const Bar<double>& bd = G.bars_d[d];
    WaveSetup W;
    wave_setup(A, d, tile, lane, W);
    if (!(bd.mut > (double)0.01f)) {  // PESysMatGen.cu:1116, :1194
        if (W.valid) out[W.v] = 0.0f;
        return;
    }
    int nb, nh;
    if (!build_candidates<kExact>(A, d, lane, W, L, nb, nh)) {
        // too many candidates for the LDS lists: report, write NaN
        if (lane == 0) atomicAdd(&A.counters[1], 1ull);
        if (W.valid) out[W.v] = __builtin_nanf("");
        return;
    }
    __syncthreads();

    using T = typename std::conditional<kExact, double, float>::type;
    KahanSum<T> acc;
    unsigned fallbacks = 0;
    for (int c = 0; c < A.sub[2]; ++c)
        for (int a = 0; a < A.sub[0]; ++a)
            for (int b = 0; b < A.sub[1]; ++b) {
                CellCtx cc;
                make_cell_ctx(bd, a, b, c, A.sub[0], A.sub[1], A.sub[2], cc);
                int nsb, nsh;
                narrow_candidates(A, cc, W, lane, L, nb, nh, nsb, nsh);
                __syncthreads();
                if (W.valid) {
                    LdsCands C{&L, nsb, nsh, lane};
                    bool near = false;
                    T val = pe_cell_ray<T>(G, d, W.p, W.ps, cc, C, near);
                    if constexpr (!kExact) {
                        if (near) {
                            bool n2 = false;
                            val = (T)pe_cell_ray<double>(G, d, W.p, W.ps, cc, C, n2);
                            ++fallbacks;
                        }
                    }
                    acc.add(val);
                }
                __syncthreads();
            }
    if (W.valid) out[W.v] = float(acc.value());
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
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
// stored.  One wave handles detector d and tiles g, g + G, ...; partial sums
// go to a per-(d, g) row of the scratch buffer (bar mu_total partials dense
// over all bars, then the layer, mu_pe and loss slots), added in a fixed
// order, and pe_grad_reduce sums the rows in a fixed order: the result does
// not depend on scheduling.

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

template <bool kExact>
__global__ __launch_bounds__(64) void pe_backward_kernel(PeLaunch A, PeGradScratch S) {
    __shared__ WaveLists L;
    __shared__ double cand_acc[kCapBars];
    extern __shared__ double row[];  // n_det partials of the bars' mu_total
    const PeGeom& G = A.geom;
    const GeomDesc& g = G.g;
    const int lane = threadIdx.x;
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    const int d = (int)(blockIdx.x / S.groups);
    const int grp = (int)(blockIdx.x % S.groups);
    const Bar<double>& bd = G.bars_d[d];
    double* out_row = S.rows + (int64_t)blockIdx.x * S.stride;
    for (int k = lane; k < g.n_det; k += 64) row[k] = 0.0;
    double lay[kMaxLayers];
    for (int m = 0; m < kMaxLayers; ++m) lay[m] = 0.0;
    double own_mut = 0.0, own_mupe = 0.0, loss = 0.0;
    unsigned fallbacks = 0;
    bool overflow = false;
    __syncthreads();
    if (bd.mut > (double)0.01f) {
        using T = typename std::conditional<kExact, double, float>::type;
        const double mut = bd.mut, mupe = bd.mupe;
        for (int tile = grp; tile < n_tiles; tile += S.groups) {
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
            for (int c = 0; c < A.sub[2]; ++c)
                for (int a = 0; a < A.sub[0]; ++a)
                    for (int b = 0; b < A.sub[1]; ++b) {
                        CellCtx cc;
                        make_cell_ctx(bd, a, b, c, A.sub[0], A.sub[1], A.sub[2], cc);
                        int nsb, nsh;
                        narrow_candidates(A, cc, W, lane, L, nb, nh, nsb, nsh);
                        __syncthreads();
                        LdsCands C{&L, nsb, nsh, lane};
                        double term = 0.0, dlay[kMaxLayers], aux[4] = {0, 0, 0, 0};
                        for (int m = 0; m < kMaxLayers; ++m) dlay[m] = 0.0;
                        bool use64 = kExact;
                        if (W.valid) {
                            bool near = false;
                            term = (double)pe_cell_ray<T>(G, d, W.p, W.ps, cc, C, near, dlay, nullptr, aux);
                            if constexpr (!kExact) {
                                if (near) {
                                    bool n2 = false;
                                    term = pe_cell_ray<double>(G, d, W.p, W.ps, cc, C, n2, dlay, nullptr, aux);
                                    use64 = true;
                                    ++fallbacks;
                                }
                            }
                        }
                        const double cw = wv * term;
                        loss += cw;
                        for (int m = 0; m < kMaxLayers; ++m) lay[m] -= cw * dlay[m];
                        const double la = aux[2], ea = aux[3], effom = aux[0];
                        own_mut += -cw * aux[1] + wv * effom * mupe * (la * ea / mut - (1.0 - ea) / (mut * mut));
                        own_mupe += wv * effom * (1.0 - ea) / mut;
                        // chords through the other bars, recomputed
                        double D[3], dd;
                        SegDir<T> sw;
                        T tol_d;
                        bool nd = false;
                        ray_dir<T>(W.p, cc.q, D, dd, sw, tol_d, nd);
                        SegDir<double> sw64;
                        double tol64;
                        ray_dir<double>(W.p, cc.q, D, dd, sw64, tol64, nd);
                        for (int j = 0; j < nsb; ++j) {
                            double x = 0.0;
                            if (W.valid && C.bar_on(j) && cw != 0.0) {
                                bool n3 = false;
                                int k = C.bar(j);
                                double l = use64 ? other_bar_chord<double>(G, k, W.p, D, dd, cc.q, sw64, tol64, n3)
                                                 : (double)other_bar_chord<T>(G, k, W.p, D, dd, cc.q, sw, tol_d, n3);
                                x = -cw * l;
                            }
                            x = wave_sum(x);
                            if (lane == 0) cand_acc[L.bar_sel[j]] += x;
                        }
                        __syncthreads();
                    }
            // fold this tile's candidate partials into the dense row (one
            // entry per candidate, each bar at most once per tile)
            for (int j = lane; j < nb; j += 64) row[L.bar_idx[j]] += cand_acc[j];
            __syncthreads();
        }
    }
    for (int m = 0; m < kMaxLayers; ++m) lay[m] = wave_sum(lay[m]);
    own_mut = wave_sum(own_mut);
    own_mupe = wave_sum(own_mupe);
    loss = wave_sum(loss);
    for (int k = lane; k < g.n_det; k += 64) out_row[k] = row[k];
    if (lane == 0) {
        out_row[d] += own_mut;
        for (int m = 0; m < g.n_layers; ++m) out_row[g.n_det + m] = lay[m];
        out_row[g.n_det + g.n_layers] = own_mupe;
        out_row[g.n_det + g.n_layers + 1] = loss;
        if (overflow) atomicAdd(&A.counters[1], 1ull);
    }
    if constexpr (!kExact) {
        unsigned tot = fallbacks;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(&A.counters[0], (unsigned long long)tot);
    }
}

// grad[0 .. Lc-1]: layers; grad[Lc + k]: mu_total of bar k; grad[Lc + N + k]:
// mu_pe of bar k; grad[Lc + 2N]: the loss.  One thread per entry, rows in order.
__global__ __launch_bounds__(256) void pe_grad_reduce(PeGradScratch S, int n_det, int n_layers,
                                                      double* grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_out = n_layers + 2 * n_det + 1;
    if (i >= n_out) return;
    double s = 0.0;
    if (i < n_layers) {
        for (int64_t r = 0; r < (int64_t)n_det * S.groups; ++r) s += S.rows[r * S.stride + n_det + i];
    } else if (i < n_layers + n_det) {
        const int k = i - n_layers;
        for (int64_t r = 0; r < (int64_t)n_det * S.groups; ++r) s += S.rows[r * S.stride + k];
    } else if (i < n_layers + 2 * n_det) {
        const int k = i - n_layers - n_det;
        for (int grp = 0; grp < S.groups; ++grp)
            s += S.rows[((int64_t)k * S.groups + grp) * S.stride + n_det + n_layers];
    } else {
        for (int64_t r = 0; r < (int64_t)n_det * S.groups; ++r)
            s += S.rows[r * S.stride + n_det + n_layers + 1];
    }
    grad[i] = s;
}

}  // namespace

hipError_t launch_pe_backward(const PeLaunch& A, const PeGradScratch& S, bool exact, double* grad,
                              hipStream_t stream) {
    const int64_t blocks = (int64_t)A.geom.g.n_det * S.groups;
    if (blocks <= 0) return hipSuccess;
    const size_t lds = sizeof(double) * (size_t)A.geom.g.n_det;
    if (exact)
        hipLaunchKernelGGL(pe_backward_kernel<true>, dim3((unsigned)blocks), dim3(64), lds, stream, A, S);
    else
        hipLaunchKernelGGL(pe_backward_kernel<false>, dim3((unsigned)blocks), dim3(64), lds, stream, A, S);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n_out = A.geom.g.n_layers + 2 * A.geom.g.n_det + 1;
    hipLaunchKernelGGL(pe_grad_reduce, dim3((n_out + 255) / 256), dim3(256), 0, stream, S,
                       A.geom.g.n_det, A.geom.g.n_layers, grad);
    return hipGetLastError();
}

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
