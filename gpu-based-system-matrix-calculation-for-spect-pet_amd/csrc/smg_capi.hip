// C ABI of the system-matrix generator: context, geometry upload, launches.
//
// Replaces the per-call device setup of the reference launch wrappers
// (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1368-1437,
// ScatterGen_RayTracing_CircularHole/scatter.cu:2013-2378): a context keeps the
// parsed geometry, the bitmaps and staging buffers on its GPU across calls
// (rotations, calibration steps), and reports errors as status codes instead of
// exiting the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/smg.h"
#include "smg_kernels.hpp"
#include "smg_parse.hpp"

using namespace smg;

namespace {

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t count) {
        if (count <= n && ptr) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&ptr, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
};

}  // namespace

// A PE call with a host output of at least 128 MiB computes the rows in up to
// this many chunks and copies each chunk to the host while the next is computed.
constexpr int kCopyChunks = 8;

// Scratch of smg_pe_forward_noalloc, the entry point for callers that cannot
// allocate: the launch counters (first 256 bytes) and the sub-cell table.
// Each GPU holds its own copy for the life of the process, shared by every
// context on that GPU.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocCounterBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct smg_ctx {
    int device = 0;
    unsigned char* scratch_dev = nullptr;  // this device's address of `scratch`
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // device-to-host copies that overlap the kernels
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_rows[kCopyChunks] = {};  // end of each row chunk of a host-output PE call
    std::string err;
    bool have_geom = false;
    HostGeom hg;
    int pe_sub[3] = {8, 16, 8};
    int sc_sub[3] = {1, 1, 1};
    DevBuf<Bar<float>> bars_f;
    DevBuf<Bar<double>> bars_d;
    DevBuf<Hole<float>> holes_f;
    DevBuf<Hole<double>> holes_d;
    DevBuf<Layer<float>> layers_f;
    DevBuf<Layer<double>> layers_d;
    DevBuf<GroupBox> bar_groups, hole_groups;
    DevBuf<float> out;        // staging for host outputs
    DevBuf<float> in;         // staging for host inputs (PE weights)
    DevBuf<float> out2, in2;  // second output and input of smg_scatter_sum
    DevBuf<unsigned long long> counters;
    DevBuf<uint32_t> bm_crystal, bm_coll;
    bool have_bm_crystal = false, have_bm_coll = false;
    DevBuf<double> scratch;   // adjoint partials
    DevBuf<double> tan_v;     // tangent direction
    DevBuf<double> grad;
    DevBuf<Sphere<float>> sph_f;       // bars, raw frame (bitmaps)
    DevBuf<Sphere<double>> sph_d;
    DevBuf<Sphere<float>> hsph_f;      // layer-0 hole centres, raw frame
    DevBuf<Sphere<double>> hsph_d;
    DevBuf<ScatterPt<double>> pts_bar;   // bar centres (crystal scatter)
    DevBuf<ScatterPt<double>> pts_hole;  // layer-0 hole centres (collimator scatter)
    DevBuf<PairTerm> pair;
    DevBuf<double> sc_partial;   // per-slice scatter sums
    DevBuf<double> sc_row_cut;   // energy-window cut per output row of a scatter launch
    DevBuf<double> proj;         // projection partial sums
    DevBuf<double> kn_tab;       // KN prefix table of kn_tab_e0
    double kn_tab_e0 = -1.0;
    DevBuf<CellCtx> cell_tab;    // PE sub-cell table of every bar
    DevBuf<CellCtx> cell_tab_pts;  // sub-cell table of one chunk of crystal grid points
    int cell_tab_sub[3] = {0, 0, 0};
    uint64_t cell_tab_gen = 0, geom_gen = 0;
    bool elliptic_holes = false;  // smg_set_hole_shapes made a hole elliptical
    MuTables mu_tab{};                 // host copy; n = 0: power laws
    DevBuf<MuTables> mu_tab_d;
    MuTablesF mu_tabf{};               // their FP32 form for the current E0
    DevBuf<MuTablesF> mu_tabf_d;
    DevBuf<ScatterPt<double>> pts_grid;  // grid scatterers (one chunk)
    DevBuf<float> w1;                    // their first-leg weights
    DevBuf<int> rows;                    // detector rows of a row-subset call
    DevBuf<double> use_rows;             // crystal grid: the rows' window data (point_used)
    DevBuf<int> spill_queue;             // PE-family work items that overflowed LDS
    DevBuf<unsigned char> spill_slots;   // their candidate lists in global memory
};

namespace {

smg_status fail(smg_ctx* c, smg_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return s;
}

#define SMG_HIP(ctx, call)                                                                \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ctx, SMG_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

template <class T>
smg_status upload(smg_ctx* c, DevBuf<T>& buf, const std::vector<T>& v) {
    SMG_HIP(c, buf.reserve(v.size()));
    if (!v.empty())
        SMG_HIP(c, hipMemcpyAsync(buf.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice,
                                  c->stream));
    return SMG_OK;
}

PeGeom device_geom(smg_ctx* c) {
    PeGeom G;
    G.g = c->hg.g;
    G.bars_f = c->bars_f.ptr;
    G.bars_d = c->bars_d.ptr;
    G.holes_f = c->holes_f.ptr;
    G.holes_d = c->holes_d.ptr;
    G.layers_f = c->layers_f.ptr;
    G.layers_d = c->layers_d.ptr;
    G.bar_groups = c->bar_groups.ptr;
    G.hole_groups = c->hole_groups.ptr;
    return G;
}

void pick_tile(const GeomDesc& g, int tile[3], int tiles[3]) {
    if (g.nz >= 4) {
        tile[0] = 4;
        tile[1] = 4;
        tile[2] = 4;
    } else if (g.nz >= 2) {
        tile[0] = 8;
        tile[1] = 4;
        tile[2] = 2;
    } else {
        tile[0] = 8;
        tile[1] = 8;
        tile[2] = 1;
    }
    tiles[0] = (g.nx + tile[0] - 1) / tile[0];
    tiles[1] = (g.ny + tile[1] - 1) / tile[1];
    tiles[2] = (g.nz + tile[2] - 1) / tile[2];
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// The point-weight, tangent and adjoint kernels run in FP64 when asked, and
// whenever a hole is elliptical: of their kernels only the all-FP64 ones
// carry the elliptic-cylinder chord.  The matrix and projection launches have
// FP32 kernels with the chord (PeLaunch::ell) and use pe_fp64.
bool pe_exact(const smg_ctx* c, smg_precision prec) {
    return prec == SMG_PREC_FP64 || c->elliptic_holes;
}
bool pe_fp64(smg_precision prec) { return prec == SMG_PREC_FP64; }

// Spill area of the PE-family launches (PeSpill): one queue entry per block
// of the largest launch on this geometry (n_det x tiles blocks for the matrix
// and the projections, at most 32,768 more for point weights and 8,192 more
// for the adjoint) and kSpillWaves list slots sized for every bar and hole.
smg_status ensure_spill(smg_ctx* c, PeLaunch& A) {
    const GeomDesc& g = c->hg.g;
    const int64_t n_tiles = (int64_t)A.tiles[0] * A.tiles[1] * A.tiles[2];
    PeSpill& S = A.spill;
    S.queue_cap = (int64_t)g.n_det * n_tiles + 65536;
    S.cap_bars = g.n_det;
    S.cap_holes = g.n_holes;
    S.slot_bytes = pe_spill_slot_bytes(g.n_det, g.n_holes);
    S.n_slots = kSpillWaves;
    SMG_HIP(c, c->spill_queue.reserve((size_t)S.queue_cap));
    SMG_HIP(c, c->spill_slots.reserve((size_t)(S.slot_bytes * S.n_slots)));
    S.queue = c->spill_queue.ptr;
    S.slots = c->spill_slots.ptr;
    return SMG_OK;
}

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress): host code uses this address, never the
// symbol itself, which on the host is only a stand-in.
hipError_t scratch_address(int device, unsigned char** ptr) {
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = (unsigned char*)p;
    *ptr = (unsigned char*)p;
    return hipSuccess;
}

// Clears n 64-bit counters (a kernel rather than a memset node, so the
// no-alloc entry point's work can be recorded into a HIP graph).
__global__ void clear_counters_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

smg_status check_ready(smg_ctx* c, int rotation) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (!c->have_geom) return fail(c, SMG_ERR_STATE, "no geometry set");
    if (rotation < 0) return fail(c, SMG_ERR_BAD_PARAM, "negative rotation index %d", rotation);
    SMG_HIP(c, hipSetDevice(c->device));
    return SMG_OK;
}


// Sum of KN(theta) sin(theta) * step over [0, pi] with the reference's float
// constants (ScatterGen_RayTracing_CircularHole/scatter.cu:1169, :980-993).
double kn_norm(double e0) {
    const double step = (double)0.01f, hi = (double)(float)3.14159265358979323846;
    int n = (int)(hi / step);
    double s = 0.0, alpha = e0 / 511.0;
    for (int i = 0; i <= n; ++i) {
        double th = i * step;
        double c = std::cos(th);
        double f1 = alpha * (1.0 - c), f2 = 1.0 + c * c;
        s += f2 / (1.0 + f1) / (1.0 + f1) * (1.0 + f1 * f1 / f2 / (1.0 + f1)) * std::sin(th) * step;
    }
    return s;
}

smg_status ensure_bitmap(smg_ctx* c, bool collimator) {
    const HostGeom& hg = c->hg;
    const int N = hg.g.n_det;
    if (collimator ? c->have_bm_coll : c->have_bm_crystal) return SMG_OK;
    int64_t n_words = collimator ? smg_bitmap_collimator_words(hg.n_holes_l0, N)
                                 : smg_bitmap_crystal_words(N);
    DevBuf<uint32_t>& buf = collimator ? c->bm_coll : c->bm_crystal;
    SMG_HIP(c, buf.reserve((size_t)n_words));
    BitmapLaunch B;
    B.spheres_f = c->sph_f.ptr;
    B.spheres_d = c->sph_d.ptr;
    B.holes_f = c->hsph_f.ptr;
    B.holes_d = c->hsph_d.ptr;
    B.words = buf.ptr;
    B.n_words = n_words;
    B.n_det = N;
    B.n_holes = hg.n_holes_l0;
    SMG_HIP(c, launch_bitmap(B, collimator, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    (collimator ? c->have_bm_coll : c->have_bm_crystal) = true;
    return SMG_OK;
}

smg_status bitmap_entry(smg_ctx* c, bool collimator, uint32_t* words, int64_t n_words,
                        int words_is_device) {
    smg_status s = check_ready(c, 0);
    if (s != SMG_OK) return s;
    const HostGeom& hg = c->hg;
    int64_t need = collimator ? smg_bitmap_collimator_words(hg.n_holes_l0, hg.g.n_det)
                              : smg_bitmap_crystal_words(hg.g.n_det);
    if (words && n_words < need)
        return fail(c, SMG_ERR_BAD_PARAM, "bitmap buffer holds %lld words, %lld needed",
                    (long long)n_words, (long long)need);
    if (collimator)
        c->have_bm_coll = false;
    else
        c->have_bm_crystal = false;
    if ((s = ensure_bitmap(c, collimator)) != SMG_OK) return s;
    if (words && need > 0) {
        const DevBuf<uint32_t>& buf = collimator ? c->bm_coll : c->bm_crystal;
        SMG_HIP(c, hipMemcpyAsync(words, buf.ptr, need * sizeof(uint32_t),
                                  words_is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                  c->stream));
        SMG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return SMG_OK;
}

smg_status load_bitmap(smg_ctx* c, bool collimator, const uint32_t* words, int64_t n_words) {
    smg_status s = check_ready(c, 0);
    if (s != SMG_OK) return s;
    const HostGeom& hg = c->hg;
    int64_t need = collimator ? smg_bitmap_collimator_words(hg.n_holes_l0, hg.g.n_det)
                              : smg_bitmap_crystal_words(hg.g.n_det);
    if (!words || n_words < need)
        return fail(c, SMG_ERR_BAD_PARAM, "bitmap holds %lld words, %lld needed",
                    (long long)n_words, (long long)need);
    DevBuf<uint32_t>& buf = collimator ? c->bm_coll : c->bm_crystal;
    SMG_HIP(c, buf.reserve((size_t)need));
    SMG_HIP(c, hipMemcpyAsync(buf.ptr, words, need * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    (collimator ? c->have_bm_coll : c->have_bm_crystal) = true;
    return SMG_OK;
}

// KN prefix table for the FP32 scatter term at the launch's E0 (built on the
// context's stream when E0 changes).
hipError_t kn_table(smg_ctx* c, ScatterLaunch& A) {
    if (c->kn_tab_e0 != A.e0) {
        c->kn_tab_e0 = -1.0;
        hipError_t e = c->kn_tab.reserve((size_t)(kKnPhase + 1) * kKnSteps);
        if (e != hipSuccess) return e;
        e = launch_kn_table(A.e0, c->kn_tab.ptr, c->stream);
        if (e != hipSuccess) return e;
        c->kn_tab_e0 = A.e0;
    }
    A.kn_table = c->kn_tab.ptr;
    return hipSuccess;
}

// Energy-window cut of each output row of a scatter launch (A.row_cut), on
// the context's stream ahead of the pair terms and scatter kernels.
hipError_t scatter_row_cuts(smg_ctx* c, ScatterLaunch& A) {
    const size_t rows = (size_t)(A.rows ? A.n_rows : A.geom.g.n_det);
    hipError_t e = c->sc_row_cut.reserve(rows > 0 ? rows : 1);
    if (e != hipSuccess) return e;
    A.row_cut = c->sc_row_cut.ptr;
    return launch_row_cuts(A, c->sc_row_cut.ptr, c->stream);
}

// Partial-sum buffer of a scatter launch whose scatter points are split over
// several waves per element.  It only grows; hipFree waits for the device.
hipError_t scatter_partials(smg_ctx* c, ScatterLaunch& A) {
    const size_t np = scatter_partial_size(A);
    if (np) {
        hipError_t e = c->sc_partial.reserve(np);
        if (e != hipSuccess) return e;
    }
    A.partial = c->sc_partial.ptr;
    return hipSuccess;
}

// Sub-cell table of every bar for the PE sub-cells of A (built on the
// context's stream when the geometry or the sub-cell counts change), or none
// when it would pass 1 GiB: the kernels then form each sub-cell themselves.
hipError_t pe_cell_table(smg_ctx* c, PeLaunch& A) {
    A.cells = nullptr;
    A.cells_bar0 = 0;
    const int64_t n = (int64_t)c->hg.g.n_det * A.sub[0] * A.sub[1] * A.sub[2];
    if (n * (int64_t)sizeof(CellCtx) > ((int64_t)1 << 30)) return hipSuccess;
    const bool same = c->cell_tab_gen == c->geom_gen && c->cell_tab_sub[0] == A.sub[0] &&
                      c->cell_tab_sub[1] == A.sub[1] && c->cell_tab_sub[2] == A.sub[2];
    if (!same) {
        c->cell_tab_gen = 0;
        hipError_t e = c->cell_tab.reserve((size_t)n);
        if (e != hipSuccess) return e;
        e = launch_cell_table(A.geom, 0, c->hg.g.n_det, A.sub, c->cell_tab.ptr, c->stream);
        if (e != hipSuccess) return e;
        c->cell_tab_gen = c->geom_gen;
        for (int k = 0; k < 3; ++k) c->cell_tab_sub[k] = A.sub[k];
    }
    A.cells = c->cell_tab.ptr;
    return hipSuccess;
}

// The mu(E) tables for photopeak e0 on the device: the tables with their
// values at e0, and their FP32 form.
hipError_t upload_mu_tables(smg_ctx* c, double e0) {
    if (c->mu_tab.n == 0) return hipSuccess;
    const double le0 = std::log(e0);
    for (int t = 0; t < 4; ++t) c->mu_tab.lmu0[t] = mu_table_log(c->mu_tab, t, le0);
    make_mu_tables_f(c->mu_tab, e0, c->mu_tabf);
    hipError_t e = c->mu_tab_d.reserve(1);
    if (e == hipSuccess) e = c->mu_tabf_d.reserve(1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->mu_tab_d.ptr, &c->mu_tab, sizeof(MuTables), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->mu_tabf_d.ptr, &c->mu_tabf, sizeof(MuTablesF), hipMemcpyHostToDevice,
                           c->stream);
    return e;
}

// Per-call set-up of the bar-centre and hole-centre scatter launches, before
// their timed region: the KN prefix table of the photopeak energy (built when
// it changes) and the mu(E) tables with their values at E0.
smg_status scatter_prep(smg_ctx* c) {
    ScatterLaunch A{};
    A.e0 = c->hg.phys[7];
    SMG_HIP(c, kn_table(c, A));
    SMG_HIP(c, upload_mu_tables(c, A.e0));
    return SMG_OK;
}

// Enqueues one scatter matrix (crystal scatter driven by the device PE block
// pe, or collimator scatter of `variant`) into the device block out, added to
// it when accumulate; no synchronisation.  Adds its terms to *rays.
smg_status scatter_launch(smg_ctx* c, bool collimator, int rotation, int variant, smg_precision prec,
                          const float* pe, float* out, bool accumulate, uint64_t* rays) {
    const HostGeom& hg = c->hg;
    const GeomDesc& g = hg.g;
    smg_status s;
    if ((s = ensure_bitmap(c, collimator)) != SMG_OK) return s;
    ScatterLaunch A{};
    A.geom = device_geom(c);
    A.kind = collimator ? 1 : 0;
    A.variant = variant;
    A.pts_d = collimator ? c->pts_hole.ptr : c->pts_bar.ptr;
    A.n_pts = collimator ? hg.n_holes_l0 : g.n_det;
    A.bitmap = collimator ? c->bm_coll.ptr : c->bm_crystal.ptr;
    A.pe = collimator ? nullptr : pe;
    const int ncell = c->sc_sub[0] * c->sc_sub[1] * c->sc_sub[2];
    SMG_HIP(c, c->pair.reserve((size_t)g.n_det * (size_t)(A.n_pts > 0 ? A.n_pts : 1) * ncell));
    A.pair = c->pair.ptr;
    A.dst = out;
    A.counters = c->counters.ptr;
    double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->sc_sub[k];
    const float* ph = hg.phys.data();
    A.compton_on = (int)std::floor((double)ph[0] + 0.5) == 1;
    A.same_window = (int)std::floor((double)ph[4] + 0.5) > 0;
    A.win_lo = ph[5];
    A.win_hi = ph[6];
    A.e0 = ph[7];
    A.kn_norm = kn_norm(A.e0);
    A.kn_table = c->kn_tab.ptr;  // scatter_prep
    A.coll_ratio = (double)hg.coll[17] / (double)hg.coll[15];
    if (collimator && !(hg.coll[15] != 0.0f)) A.coll_ratio = 0.0;
    A.tables = c->mu_tab.n > 0 ? c->mu_tab_d.ptr : nullptr;
    A.tables_f = c->mu_tab.n > 0 ? c->mu_tabf_d.ptr : nullptr;
    A.w1 = nullptr;
    A.accumulate = accumulate ? 1 : 0;
    A.rows = nullptr;
    A.n_rows = 0;
    A.pts_per_bar = 1;
    A.pt_bar0 = 0;
    SMG_HIP(c, scatter_row_cuts(c, A));
    SMG_HIP(c, launch_pair_terms(A, c->stream));
    SMG_HIP(c, scatter_partials(c, A));
    SMG_HIP(c, launch_scatter(A, prec == SMG_PREC_FP64, c->stream));
    *rays += (uint64_t)g.n_det * (uint64_t)g.n_vox * (uint64_t)(A.n_pts > 0 ? A.n_pts : 0) * (uint64_t)ncell;
    return SMG_OK;
}

// The PE block of a scatter call on the device (uploaded into buf when it is
// on the host).
smg_status device_block(smg_ctx* c, const float* blk, int is_device, DevBuf<float>& buf,
                        const float** out) {
    if (is_device) {
        *out = blk;
        return SMG_OK;
    }
    const size_t n_el = (size_t)c->hg.g.n_det * (size_t)c->hg.g.n_vox;
    SMG_HIP(c, buf.reserve(n_el));
    SMG_HIP(c, hipMemcpyAsync(buf.ptr, blk, n_el * sizeof(float), hipMemcpyHostToDevice, c->stream));
    *out = buf.ptr;
    return SMG_OK;
}

smg_status scatter_entry(smg_ctx* c, bool collimator, int rotation, int variant,
                         smg_precision prec, const float* pe, int pe_is_device, float* dst,
                         int dst_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dst) return fail(c, SMG_ERR_BAD_PARAM, "dst is NULL");
    if (!collimator && !pe) return fail(c, SMG_ERR_BAD_PARAM, "PE block is NULL");
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    const float* pe_d = nullptr;
    if (!collimator && (s = device_block(c, pe, pe_is_device, c->in, &pe_d)) != SMG_OK) return s;
    float* out = dst;
    if (!dst_is_device) {
        SMG_HIP(c, c->out.reserve(n_el));
        out = c->out.ptr;
    }
    // the bitmap and the tables are built before the timed region
    if ((s = ensure_bitmap(c, collimator)) != SMG_OK) return s;
    if ((s = scatter_prep(c)) != SMG_OK) return s;
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    uint64_t rays = 0;
    if ((s = scatter_launch(c, collimator, rotation, variant, prec, pe_d, out, false, &rays)) != SMG_OK)
        return s;
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!dst_is_device)
        SMG_HIP(c, hipMemcpyAsync(dst, out, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = rays;
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = 0;
    }
    return SMG_OK;
}

}  // namespace

// ============================================================================
extern "C" {

const char* smg_status_string(smg_status s) {
    switch (s) {
        case SMG_OK: return "ok";
        case SMG_ERR_BAD_PARAM: return "bad parameter";
        case SMG_ERR_CAPACITY: return "capacity exceeded";
        case SMG_ERR_IO: return "i/o error";
        case SMG_ERR_HIP: return "HIP runtime error";
        case SMG_ERR_NO_DEVICE: return "no device";
        case SMG_ERR_STATE: return "invalid call order";
    }
    return "unknown status";
}

smg_status smg_device_count(int* count) {
    if (!count) return SMG_ERR_BAD_PARAM;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SMG_OK;
}

smg_status smg_create(int device, smg_ctx** out) {
    if (!out) return SMG_ERR_BAD_PARAM;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SMG_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SMG_ERR_NO_DEVICE;
    smg_ctx* c = new smg_ctx();
    c->device = device;
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreate(&c->stream) == hipSuccess &&
              hipStreamCreate(&c->copy_stream) == hipSuccess;
    for (int i = 0; ok && i < kCopyChunks; ++i)
        ok = hipEventCreateWithFlags(&c->ev_rows[i], hipEventDisableTiming) == hipSuccess;
    if (!ok || hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        c->counters.reserve(4) != hipSuccess || scratch_address(device, &c->scratch_dev) != hipSuccess) {
        smg_destroy(c);
        return SMG_ERR_HIP;
    }
    *out = c;
    return SMG_OK;
}

smg_status smg_destroy(smg_ctx* c) {
    if (!c) return SMG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->bars_f.release();
    c->bars_d.release();
    c->holes_f.release();
    c->holes_d.release();
    c->layers_f.release();
    c->layers_d.release();
    c->bar_groups.release();
    c->hole_groups.release();
    c->out.release();
    c->in.release();
    c->counters.release();
    c->bm_crystal.release();
    c->bm_coll.release();
    c->scratch.release();
    c->grad.release();
    c->sph_f.release();
    c->sph_d.release();
    c->hsph_f.release();
    c->hsph_d.release();
    c->pts_bar.release();
    c->pts_hole.release();
    c->pair.release();
    c->sc_partial.release();
    c->sc_row_cut.release();
    c->proj.release();
    c->kn_tab.release();
    c->cell_tab.release();
    c->cell_tab_pts.release();
    c->mu_tab_d.release();
    c->mu_tabf_d.release();
    c->pts_grid.release();
    c->w1.release();
    c->rows.release();
    c->out2.release();
    c->in2.release();
    c->tan_v.release();
    c->spill_queue.release();
    c->spill_slots.release();
    c->use_rows.release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t& e : c->ev_rows)
        if (e) (void)hipEventDestroy(e);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SMG_OK;
}

const char* smg_last_error(const smg_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* smg_stream(smg_ctx* c) { return c ? (void*)c->stream : nullptr; }

smg_status smg_synchronize(smg_ctx* c) {
    if (!c) return SMG_ERR_BAD_PARAM;
    SMG_HIP(c, hipSetDevice(c->device));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    return SMG_OK;
}

smg_status smg_set_geometry(smg_ctx* c, const float* coll, size_t n_coll, const float* det,
                            size_t n_det, const float* img, size_t n_img, const float* phys,
                            size_t n_phys) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (!coll || !det || !img || !phys || n_coll < 80000 || n_det < 80000 || n_img < 100 ||
        n_phys < 100)
        return fail(c, SMG_ERR_BAD_PARAM,
                    "parameter arrays must hold 80000/80000/100/100 floats (got %zu/%zu/%zu/%zu)",
                    n_coll, n_det, n_img, n_phys);
    SMG_HIP(c, hipSetDevice(c->device));
    HostGeom hg;
    std::string why;
    int code = parse_geometry(coll, det, img, phys, hg, why);
    if (code != 0) return fail(c, (smg_status)code, "%s", why.c_str());
    smg_status s;
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    // the device buffers are replaced one by one: until all of them hold the
    // new geometry the context has none, so a failed upload leaves it refusing
    // launches instead of mixing two geometries
    c->have_geom = false;
    c->have_bm_crystal = c->have_bm_coll = false;
    if ((s = upload(c, c->bars_f, hg.bars_f)) != SMG_OK) return s;
    if ((s = upload(c, c->bars_d, hg.bars_d)) != SMG_OK) return s;
    if ((s = upload(c, c->holes_f, hg.holes_f)) != SMG_OK) return s;
    if ((s = upload(c, c->holes_d, hg.holes_d)) != SMG_OK) return s;
    if ((s = upload(c, c->layers_f, hg.layers_f)) != SMG_OK) return s;
    if ((s = upload(c, c->layers_d, hg.layers_d)) != SMG_OK) return s;
    if ((s = upload(c, c->bar_groups, hg.bar_groups)) != SMG_OK) return s;
    if ((s = upload(c, c->hole_groups, hg.hole_groups)) != SMG_OK) return s;
    std::vector<Sphere<float>> sf;
    std::vector<Sphere<double>> sd, hd;
    std::vector<Sphere<float>> hf;
    std::vector<ScatterPt<double>> pb, ph;
    make_scatter_records(hg, sf, sd, hf, hd, pb, ph);
    if ((s = upload(c, c->sph_f, sf)) != SMG_OK) return s;
    if ((s = upload(c, c->sph_d, sd)) != SMG_OK) return s;
    if ((s = upload(c, c->hsph_f, hf)) != SMG_OK) return s;
    if ((s = upload(c, c->hsph_d, hd)) != SMG_OK) return s;
    if ((s = upload(c, c->pts_bar, pb)) != SMG_OK) return s;
    if ((s = upload(c, c->pts_hole, ph)) != SMG_OK) return s;
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    c->hg = std::move(hg);
    c->have_geom = true;
    c->have_bm_crystal = c->have_bm_coll = false;
    c->elliptic_holes = false;
    ++c->geom_gen;
    return SMG_OK;
}

smg_status smg_set_hole_shapes(smg_ctx* c, const float* shapes, int n_holes) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (!c->have_geom) return fail(c, SMG_ERR_STATE, "no geometry set");
    SMG_HIP(c, hipSetDevice(c->device));
    HostGeom hg = c->hg;
    std::string why;
    int code = apply_hole_shapes(hg, shapes, n_holes, why);
    if (code != 0) return fail(c, (smg_status)code, "%s", why.c_str());
    smg_status s;
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    c->have_geom = false;  // until every hole buffer holds the new shapes
    if ((s = upload(c, c->holes_f, hg.holes_f)) != SMG_OK) return s;
    if ((s = upload(c, c->holes_d, hg.holes_d)) != SMG_OK) return s;
    if ((s = upload(c, c->hole_groups, hg.hole_groups)) != SMG_OK) return s;
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    c->hg = std::move(hg);
    c->have_geom = true;
    c->elliptic_holes = false;
    for (const Hole<double>& o : c->hg.holes_d) c->elliptic_holes |= o.erx > 0.0;
    return SMG_OK;
}

smg_status smg_get_dims(const smg_ctx* c, int* n_det, int64_t* n_vox, int* n_rot, int* n_holes_l0) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (!c->have_geom) return SMG_ERR_STATE;
    if (n_det) *n_det = c->hg.g.n_det;
    if (n_vox) *n_vox = c->hg.g.n_vox;
    if (n_rot) *n_rot = c->hg.n_rot;
    if (n_holes_l0) *n_holes_l0 = c->hg.n_holes_l0;
    return SMG_OK;
}

smg_status smg_set_pe_subsamples(smg_ctx* c, int sx, int sy, int sz) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (sx < 1 || sy < 1 || sz < 1 || (int64_t)sx * sy * sz > 65536)
        return fail(c, SMG_ERR_BAD_PARAM, "sub-cells %d x %d x %d", sx, sy, sz);
    c->pe_sub[0] = sx;
    c->pe_sub[1] = sy;
    c->pe_sub[2] = sz;
    return SMG_OK;
}

smg_status smg_set_mu_tables(smg_ctx* c, int n_nodes, const double* energy_kev,
                             const double* crystal_pe, const double* crystal_compton,
                             const double* coll_pe, const double* coll_compton) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (n_nodes == 0) {
        SMG_HIP(c, hipStreamSynchronize(c->stream));
        c->mu_tab.n = 0;
        return SMG_OK;
    }
    if (n_nodes < 2 || n_nodes > kMaxMuNodes)
        return fail(c, SMG_ERR_BAD_PARAM, "mu(E) tables need 2..%d nodes (got %d)", kMaxMuNodes, n_nodes);
    if (!energy_kev || !crystal_pe || !crystal_compton || !coll_pe || !coll_compton)
        return fail(c, SMG_ERR_BAD_PARAM, "mu(E) table pointer is NULL");
    const double* v[4] = {crystal_pe, crystal_compton, coll_pe, coll_compton};
    MuTables t{};
    t.n = n_nodes;
    for (int i = 0; i < n_nodes; ++i) {
        if (!(energy_kev[i] > 0.0) || (i > 0 && !(energy_kev[i] > energy_kev[i - 1])))
            return fail(c, SMG_ERR_BAD_PARAM, "mu(E) node energies must be positive and increasing");
        t.le[i] = std::log(energy_kev[i]);
        for (int k = 0; k < 4; ++k) {
            if (!(v[k][i] > 0.0)) return fail(c, SMG_ERR_BAD_PARAM, "mu(E) values must be positive");
            t.lmu[k][i] = std::log(v[k][i]);
        }
    }
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    c->mu_tab = t;
    return SMG_OK;
}

smg_status smg_set_scatter_subsamples(smg_ctx* c, int sx, int sy, int sz) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (sx < 1 || sy < 1 || sz < 1 || (int64_t)sx * sy * sz > 4096)
        return fail(c, SMG_ERR_BAD_PARAM, "sub-cells %d x %d x %d", sx, sy, sz);
    c->sc_sub[0] = sx;
    c->sc_sub[1] = sy;
    c->sc_sub[2] = sz;
    return SMG_OK;
}

smg_status smg_pe_forward(smg_ctx* c, int rotation, smg_precision prec, float* dst,
                          int dst_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dst) return fail(c, SMG_ERR_BAD_PARAM, "dst is NULL");
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    float* out = dst;
    if (!dst_is_device) {
        SMG_HIP(c, c->out.reserve(n_el));
        out = c->out.ptr;
    }
    PeLaunch A{};
    A.geom = device_geom(c);
    A.dst = out;
    A.counters = c->counters.ptr;
    double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    pick_tile(g, A.tile, A.tiles);
    A.ell = c->elliptic_holes;
    if ((s = ensure_spill(c, A)) != SMG_OK) return s;
    SMG_HIP(c, pe_cell_table(c, A));
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    const bool exact = pe_fp64(prec);
    if (dst_is_device) {
        SMG_HIP(c, launch_pe_forward(A, exact, c->stream));
        SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    } else {
        // host output: rows in chunks of at least 64 MiB, each copied on the
        // copy stream once computed, so the copies overlap the following
        // chunks' kernels (a small block stays one launch)
        const size_t bytes = n_el * sizeof(float);
        const int n_chunks = (int)std::max<size_t>(
            1, std::min<size_t>({(size_t)kCopyChunks, (size_t)g.n_det, bytes >> 26}));
        int row0[kCopyChunks + 1];
        for (int i = 0; i <= n_chunks; ++i) row0[i] = (int)((int64_t)g.n_det * i / n_chunks);
        for (int i = 0; i < n_chunks; ++i) {
            SMG_HIP(c, launch_pe_forward_rows(A, exact, row0[i], row0[i + 1] - row0[i], c->stream));
            SMG_HIP(c, hipEventRecord(c->ev_rows[i], c->stream));
        }
        SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
        for (int i = 0; i < n_chunks; ++i) {
            const size_t off = (size_t)row0[i] * (size_t)g.n_vox;
            const size_t n = (size_t)(row0[i + 1] - row0[i]) * (size_t)g.n_vox;
            SMG_HIP(c, hipStreamWaitEvent(c->copy_stream, c->ev_rows[i], 0));
            SMG_HIP(c, hipMemcpyAsync(dst + off, out + off, n * sizeof(float), hipMemcpyDeviceToHost,
                                      c->copy_stream));
        }
    }
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    if (!dst_is_device) SMG_HIP(c, hipStreamSynchronize(c->copy_stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)(c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2]);
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    return SMG_OK;
}

// smg_pe_forward into a caller's device buffer with no allocation: the
// counters and the sub-cell table live in this device's `scratch`, and tiles
// beyond the LDS candidate lists (which smg_pe_forward redoes with lists it
// allocates) are refused.
smg_status smg_pe_forward_noalloc(smg_ctx* c, int rotation, smg_precision prec, float* dst_device,
                                  smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dst_device) return fail(c, SMG_ERR_BAD_PARAM, "dst is NULL");
    if (!c->scratch_dev) return fail(c, SMG_ERR_STATE, "no scratch address for device %d", c->device);
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    const size_t n_cells = (size_t)c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2];
    const size_t need = kNoAllocCounterBytes + (size_t)g.n_det * n_cells * sizeof(CellCtx);
    if (need > kNoAllocScratchBytes)
        return fail(c, SMG_ERR_CAPACITY,
                    "%d bars x %zu sub-cells need %zu bytes of scratch; smg_pe_forward_noalloc has %zu",
                    g.n_det, n_cells, need, kNoAllocScratchBytes);
    unsigned long long* counters = (unsigned long long*)c->scratch_dev;
    CellCtx* cells = (CellCtx*)(c->scratch_dev + kNoAllocCounterBytes);
    PeLaunch A{};
    A.geom = device_geom(c);
    A.dst = dst_device;
    A.counters = counters;
    const double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    pick_tile(g, A.tile, A.tiles);
    A.cells = cells;
    A.cells_bar0 = 0;
    A.spill = PeSpill{};  // no spill area: overflowing work items are only counted
    A.ell = c->elliptic_holes;
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, c->stream, counters, 4);
    SMG_HIP(c, hipGetLastError());
    SMG_HIP(c, launch_cell_table(A.geom, 0, g.n_det, A.sub, cells, c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    SMG_HIP(c, launch_pe_forward(A, pe_fp64(prec), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)n_cells;
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[2];
    }
    if (cnt[2])
        return fail(c, SMG_ERR_CAPACITY,
                    "%llu voxel tiles need more than %d bars or %d holes of candidates; "
                    "smg_pe_forward_noalloc has no spill area (smg_pe_forward handles them)",
                    cnt[2], kCapBars, kCapHoles);
    return SMG_OK;
}

// Matrix-free projections (smg_pe_project / smg_pe_backproject).
static smg_status project_entry(smg_ctx* c, int rotation, smg_precision prec, bool back, const float* in,
                                int in_is_device, float* out, int out_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!in || !out) return fail(c, SMG_ERR_BAD_PARAM, "input and output must not be NULL");
    const GeomDesc& g = c->hg.g;
    const size_t n_in = back ? (size_t)g.n_det : (size_t)g.n_vox;
    const size_t n_out = back ? (size_t)g.n_vox : (size_t)g.n_det;
    PeLaunch A{};
    A.geom = device_geom(c);
    A.counters = c->counters.ptr;
    const double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    A.cell_slices = 1;
    A.ell = c->elliptic_holes;
    pick_tile(g, A.tile, A.tiles);
    if ((s = ensure_spill(c, A)) != SMG_OK) return s;
    SMG_HIP(c, pe_cell_table(c, A));
    const float* in_d = in;
    if (!in_is_device) {
        SMG_HIP(c, c->in.reserve(n_in));
        SMG_HIP(c, hipMemcpyAsync(c->in.ptr, in, n_in * sizeof(float), hipMemcpyHostToDevice, c->stream));
        in_d = c->in.ptr;
    }
    float* out_d = out;
    if (!out_is_device) {
        SMG_HIP(c, c->out.reserve(n_out));
        out_d = c->out.ptr;
    }
    SMG_HIP(c, c->proj.reserve(project_scratch(A, back)));
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (back)
        SMG_HIP(c, launch_pe_backproject(A, pe_fp64(prec), in_d, c->proj.ptr, out_d, c->stream));
    else
        SMG_HIP(c, launch_pe_project(A, pe_fp64(prec), in_d, c->proj.ptr, out_d, c->stream));
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!out_is_device)
        SMG_HIP(c, hipMemcpyAsync(out, out_d, n_out * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        const uint64_t n_el = (uint64_t)g.n_det * (uint64_t)g.n_vox;
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)(c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2]);
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    return SMG_OK;
}

smg_status smg_pe_project(smg_ctx* c, int rotation, smg_precision prec, const float* x, int x_is_device,
                          float* y, int y_is_device, smg_stats* stats) {
    return project_entry(c, rotation, prec, false, x, x_is_device, y, y_is_device, stats);
}

smg_status smg_pe_backproject(smg_ctx* c, int rotation, smg_precision prec, const float* y,
                              int y_is_device, float* x, int x_is_device, smg_stats* stats) {
    return project_entry(c, rotation, prec, true, y, y_is_device, x, x_is_device, stats);
}

int64_t smg_bitmap_crystal_words(int n_det) {
    if (n_det < 0) return 0;
    int64_t bits = (int64_t)n_det * (n_det + 1) / 2 * n_det;
    return (bits + 31) / 32;
}

int64_t smg_bitmap_collimator_words(int n_holes_l0, int n_det) {
    if (n_det < 0 || n_holes_l0 < 0) return 0;
    int64_t bits = (int64_t)n_holes_l0 * n_det * n_det;
    return (bits + 31) / 32;
}

smg_status smg_bitmap_crystal(smg_ctx* c, uint32_t* words, int64_t n_words, int words_is_device) {
    return bitmap_entry(c, false, words, n_words, words_is_device);
}

smg_status smg_bitmap_collimator(smg_ctx* c, uint32_t* words, int64_t n_words, int words_is_device) {
    return bitmap_entry(c, true, words, n_words, words_is_device);
}

smg_status smg_load_bitmap_crystal(smg_ctx* c, const uint32_t* words, int64_t n_words) {
    return load_bitmap(c, false, words, n_words);
}

smg_status smg_load_bitmap_collimator(smg_ctx* c, const uint32_t* words, int64_t n_words) {
    return load_bitmap(c, true, words, n_words);
}

smg_status smg_scatter_crystal(smg_ctx* c, int rotation, smg_precision prec, const float* pe,
                               int pe_is_device, float* dst, int dst_is_device, smg_stats* stats) {
    return scatter_entry(c, false, rotation, 0, prec, pe, pe_is_device, dst, dst_is_device, stats);
}

smg_status smg_scatter_collimator(smg_ctx* c, int rotation, smg_coll_variant variant,
                                  smg_precision prec, float* dst, int dst_is_device,
                                  smg_stats* stats) {
    if (variant != SMG_COLL_COMBINED && variant != SMG_COLL_STANDALONE)
        return fail(c, SMG_ERR_BAD_PARAM, "unknown collimator variant %d", (int)variant);
    return scatter_entry(c, true, rotation, (int)variant, prec, nullptr, 0, dst, dst_is_device, stats);
}

smg_status smg_scatter_sum(smg_ctx* c, int rotation, unsigned parts, smg_precision prec,
                           const float* pe_drive, const float* pe_add, int pe_is_device,
                           float* scatter_dst, float* sum_dst, int dst_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    const unsigned known = SMG_PART_CRYSTAL | SMG_PART_COLL_COMBINED | SMG_PART_COLL_STANDALONE;
    if (parts == 0 || (parts & ~known))
        return fail(c, SMG_ERR_BAD_PARAM, "scatter parts 0x%x (crystal 1, collimator 2 or 4)", parts);
    if ((parts & SMG_PART_COLL_COMBINED) && (parts & SMG_PART_COLL_STANDALONE))
        return fail(c, SMG_ERR_BAD_PARAM, "one collimator variant per call");
    if (!scatter_dst && !sum_dst) return fail(c, SMG_ERR_BAD_PARAM, "no output");
    if ((parts & SMG_PART_CRYSTAL) && !pe_drive)
        return fail(c, SMG_ERR_BAD_PARAM, "crystal scatter needs the PE block that drives it");
    if (sum_dst && !pe_add) return fail(c, SMG_ERR_BAD_PARAM, "the sum needs the PE block to add");
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    const float *drive_d = nullptr, *add_d = nullptr;
    if ((parts & SMG_PART_CRYSTAL) && (s = device_block(c, pe_drive, pe_is_device, c->in, &drive_d)) != SMG_OK)
        return s;
    if (sum_dst) {
        if (pe_add == pe_drive && drive_d)
            add_d = drive_d;
        else if ((s = device_block(c, pe_add, pe_is_device, c->in2, &add_d)) != SMG_OK)
            return s;
    }
    float* sc = scatter_dst;
    if (!dst_is_device || !scatter_dst) {
        SMG_HIP(c, c->out.reserve(n_el));
        sc = c->out.ptr;
    }
    float* sm = sum_dst;
    if (sum_dst && !dst_is_device) {
        SMG_HIP(c, c->out2.reserve(n_el));
        sm = c->out2.ptr;
    }
    if ((parts & SMG_PART_CRYSTAL) && (s = ensure_bitmap(c, false)) != SMG_OK) return s;
    if ((parts & ~SMG_PART_CRYSTAL) && (s = ensure_bitmap(c, true)) != SMG_OK) return s;
    if ((s = scatter_prep(c)) != SMG_OK) return s;
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    uint64_t rays = 0;
    bool acc = false;
    if (parts & SMG_PART_CRYSTAL) {
        if ((s = scatter_launch(c, false, rotation, 0, prec, drive_d, sc, acc, &rays)) != SMG_OK) return s;
        acc = true;
    }
    if (parts & (SMG_PART_COLL_COMBINED | SMG_PART_COLL_STANDALONE)) {
        const int variant = (parts & SMG_PART_COLL_STANDALONE) ? 1 : 0;
        if ((s = scatter_launch(c, true, rotation, variant, prec, nullptr, sc, acc, &rays)) != SMG_OK) return s;
    }
    if (sum_dst) SMG_HIP(c, launch_add_blocks(add_d, sc, sm, n_el, c->stream));
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!dst_is_device) {
        if (scatter_dst)
            SMG_HIP(c, hipMemcpyAsync(scatter_dst, sc, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (sum_dst)
            SMG_HIP(c, hipMemcpyAsync(sum_dst, sm, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = rays;
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = 0;
    }
    return SMG_OK;
}

smg_status smg_scatter_collimator_grid(smg_ctx* c, int rotation, double pitch_mm, unsigned layer_mask,
                                       smg_precision prec, float* dst, int dst_is_device,
                                       int64_t* n_points, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dst) return fail(c, SMG_ERR_BAD_PARAM, "dst is NULL");
    if (!(pitch_mm > 0.0)) return fail(c, SMG_ERR_BAD_PARAM, "grid pitch must be positive");
    const HostGeom& hg = c->hg;
    const GeomDesc& g = hg.g;
    if (g.n_layers < 1) return fail(c, SMG_ERR_BAD_PARAM, "the collimator has no layer");
    if (layer_mask == 0 || (g.n_layers < 32 && (layer_mask >> g.n_layers) != 0))
        return fail(c, SMG_ERR_BAD_PARAM, "layer mask 0x%x does not select layers of the %d-layer collimator",
                    layer_mask, g.n_layers);
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    std::vector<ScatterPt<double>> pts;
    make_coll_grid(hg, pitch_mm, layer_mask, pts);
    if (pts.size() > (size_t)INT32_MAX)
        return fail(c, SMG_ERR_CAPACITY, "%zu grid points", pts.size());
    if (n_points) *n_points = (int64_t)pts.size();
    if ((s = upload(c, c->pts_grid, pts)) != SMG_OK) return s;
    const int ncell = c->sc_sub[0] * c->sc_sub[1] * c->sc_sub[2];
    // chunks of at most 65536 scatter points, first-leg weights <= 2^28 floats,
    // pair terms <= 2^25 records
    int64_t chunk = std::min<int64_t>(65536, std::max<int64_t>(1, ((int64_t)1 << 28) / std::max<int64_t>(1, g.n_vox)));
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, ((int64_t)1 << 25) / ((int64_t)g.n_det * ncell)));
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, (int64_t)pts.size()));
    SMG_HIP(c, c->w1.reserve((size_t)chunk * (size_t)g.n_vox));
    SMG_HIP(c, c->pair.reserve((size_t)g.n_det * (size_t)chunk * ncell));
    ScatterLaunch A{};
    A.geom = device_geom(c);
    A.kind = 2;
    A.variant = 0;
    A.bitmap = nullptr;
    A.pe = nullptr;
    A.pair = c->pair.ptr;
    A.w1 = c->w1.ptr;
    float* out = dst;
    if (!dst_is_device) {
        SMG_HIP(c, c->out.reserve(n_el));
        out = c->out.ptr;
    }
    A.dst = out;
    A.counters = c->counters.ptr;
    const double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->sc_sub[k];
    const float* ph = hg.phys.data();
    A.compton_on = (int)std::floor((double)ph[0] + 0.5) == 1;
    A.same_window = (int)std::floor((double)ph[4] + 0.5) > 0;
    A.win_lo = ph[5];
    A.win_hi = ph[6];
    A.e0 = ph[7];
    A.kn_norm = kn_norm(A.e0);
    SMG_HIP(c, kn_table(c, A));
    A.coll_ratio = 0.0;
    A.tables = nullptr;
    A.tables_f = nullptr;
    if (c->mu_tab.n > 0) {
        SMG_HIP(c, upload_mu_tables(c, A.e0));
        A.tables = c->mu_tab_d.ptr;
        A.tables_f = c->mu_tabf_d.ptr;
    }
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (pts.empty()) {
        SMG_HIP(c, hipMemsetAsync(out, 0, n_el * sizeof(float), c->stream));
    }
    for (int64_t s0 = 0; s0 < (int64_t)pts.size(); s0 += chunk) {
        A.pts_d = c->pts_grid.ptr + s0;
        A.n_pts = (int)std::min<int64_t>(chunk, (int64_t)pts.size() - s0);
        A.accumulate = s0 > 0;
        SMG_HIP(c, launch_first_leg(A, c->w1.ptr, c->stream));
        SMG_HIP(c, scatter_row_cuts(c, A));
        SMG_HIP(c, launch_pair_terms(A, c->stream));
        SMG_HIP(c, scatter_partials(c, A));
        SMG_HIP(c, launch_scatter(A, prec == SMG_PREC_FP64, c->stream));
    }
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!dst_is_device)
        SMG_HIP(c, hipMemcpyAsync(dst, out, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)pts.size() * (uint64_t)ncell;
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = 0;
    }
    return SMG_OK;
}

smg_status smg_scatter_crystal_grid(smg_ctx* c, int rotation, int gx, int gy, int gz,
                                    const int* rows, int n_rows, smg_precision prec, float* dst,
                                    int dst_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dst) return fail(c, SMG_ERR_BAD_PARAM, "dst is NULL");
    const HostGeom& hg = c->hg;
    const GeomDesc& g = hg.g;
    if (gx < 1 || gy < 1 || gz < 1 || (int64_t)gx * gy * gz > 65536)
        return fail(c, SMG_ERR_BAD_PARAM, "grid cells %d x %d x %d", gx, gy, gz);
    const int npb = gx * gy * gz;
    std::vector<int> row_list;
    if (rows) {
        if (n_rows < 1) return fail(c, SMG_ERR_BAD_PARAM, "n_rows must be positive");
        for (int i = 0; i < n_rows; ++i) {
            if (rows[i] < 0 || rows[i] >= g.n_det)
                return fail(c, SMG_ERR_BAD_PARAM, "row %d: bar %d out of range", i, rows[i]);
            row_list.push_back(rows[i]);
        }
    } else {
        n_rows = g.n_det;
    }
    const size_t n_el = (size_t)n_rows * (size_t)g.n_vox;
    if ((s = ensure_bitmap(c, false)) != SMG_OK) return s;
    if (rows && (s = upload(c, c->rows, row_list)) != SMG_OK) return s;
    const int ncell = c->sc_sub[0] * c->sc_sub[1] * c->sc_sub[2];
    // chunks of bars: first-leg weights <= 2^28 floats, pair terms <= 2^25 records
    int64_t nbc = std::max<int64_t>(1, ((int64_t)1 << 28) / ((int64_t)npb * g.n_vox));
    nbc = std::min<int64_t>(nbc, std::max<int64_t>(1, ((int64_t)1 << 25) / ((int64_t)n_rows * npb * ncell)));
    nbc = std::min<int64_t>(nbc, g.n_det);
    SMG_HIP(c, c->w1.reserve((size_t)nbc * npb * (size_t)g.n_vox));
    SMG_HIP(c, c->pair.reserve((size_t)n_rows * (size_t)nbc * npb * ncell));
    SMG_HIP(c, c->pts_grid.reserve((size_t)nbc * npb));
    float* out = dst;
    if (!dst_is_device) {
        SMG_HIP(c, c->out.reserve(n_el));
        out = c->out.ptr;
    }
    const double phi = (double)rotation * g.dphi;
    PeLaunch P{};
    P.geom = device_geom(c);
    P.dst = nullptr;
    P.counters = c->counters.ptr;
    P.cphi = std::cos(phi);
    P.sphi = std::sin(phi);
    P.sub[0] = gx;
    P.sub[1] = gy;
    P.sub[2] = gz;
    pick_tile(g, P.tile, P.tiles);
    if ((s = ensure_spill(c, P)) != SMG_OK) return s;
    P.pts = c->w1.ptr;
    // sub-cell table of each chunk's bars (grid cells), up to 1 GiB
    const bool cell_table = (int64_t)nbc * npb * (int64_t)sizeof(CellCtx) <= ((int64_t)1 << 30);
    if (cell_table) SMG_HIP(c, c->cell_tab_pts.reserve((size_t)nbc * npb));
    ScatterLaunch A{};
    A.geom = P.geom;
    A.kind = 3;
    A.bitmap = c->bm_crystal.ptr;
    A.pair = c->pair.ptr;
    A.w1 = c->w1.ptr;
    A.pts_d = c->pts_grid.ptr;
    A.dst = out;
    A.counters = c->counters.ptr;
    A.cphi = P.cphi;
    A.sphi = P.sphi;
    for (int k = 0; k < 3; ++k) A.sub[k] = c->sc_sub[k];
    const float* ph = hg.phys.data();
    A.compton_on = (int)std::floor((double)ph[0] + 0.5) == 1;
    A.same_window = (int)std::floor((double)ph[4] + 0.5) > 0;
    A.win_lo = ph[5];
    A.win_hi = ph[6];
    A.e0 = ph[7];
    A.kn_norm = kn_norm(A.e0);
    SMG_HIP(c, kn_table(c, A));
    A.tables = nullptr;
    A.tables_f = nullptr;
    if (c->mu_tab.n > 0) {
        SMG_HIP(c, upload_mu_tables(c, A.e0));
        A.tables = c->mu_tab_d.ptr;
        A.tables_f = c->mu_tabf_d.ptr;
    }
    A.rows = rows ? c->rows.ptr : nullptr;
    A.n_rows = n_rows;
    A.pts_per_bar = npb;
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    // rows that will read the point weights: bar centre, bounding radius and
    // window cosine cut each, so the points kernel can skip weights no row
    // uses (one lane per row: up to 64 rows)
    P.use_rows = nullptr;
    P.n_use_rows = 0;
    {
        const int nr = rows ? n_rows : g.n_det;
        if (nr <= 64) {
            std::vector<double> ur((size_t)nr * 5);
            for (int i = 0; i < nr; ++i) {
                const int d = rows ? rows[i] : i;
                const Bar<double>& bd = hg.bars_d[d];
                ScatterPhys sp;
                sp.e0 = A.e0;
                sp.lo = A.same_window ? A.win_lo : (1 - bd.res / 2.0) * A.e0;
                sp.hi = A.same_window ? A.win_hi : (1 + bd.res / 2.0) * A.e0;
                sp.res = bd.res;
                ur[5 * i + 0] = bd.cx;
                ur[5 * i + 1] = bd.cy;
                ur[5 * i + 2] = bd.cz;
                ur[5 * i + 3] = 0.5 * std::sqrt(bd.w * bd.w + bd.t * bd.t + bd.h * bd.h) * (1.0 + 1e-9);
                ur[5 * i + 4] = window_reject_cos(sp, 0.0);
            }
            if ((s = upload(c, c->use_rows, ur)) != SMG_OK) return s;
            P.use_rows = c->use_rows.ptr;
            P.n_use_rows = nr;
        }
    }
    std::vector<ScatterPt<double>> pts((size_t)nbc * npb);
    const int sub[3] = {gx, gy, gz};
    for (int b0 = 0; b0 < g.n_det; b0 += (int)nbc) {
        const int nb = (int)std::min<int64_t>(nbc, g.n_det - b0);
        // grid points: the sub-cell centres, in the PE kernel's sub-cell order
        SMG_HIP(c, hipStreamSynchronize(c->stream));  // pts is reused
        for (int b = 0; b < nb; ++b)
            for (int cell = 0; cell < npb; ++cell) {
                CellCtx cc;
                make_cell_ctx_linear(hg.bars_d[b0 + b], cell, sub, cc);
                pts[(size_t)b * npb + cell] = ScatterPt<double>{cc.q[0], cc.q[1], cc.q[2], 0.0};
            }
        SMG_HIP(c, hipMemcpyAsync(c->pts_grid.ptr, pts.data(), sizeof(ScatterPt<double>) * nb * npb,
                                  hipMemcpyHostToDevice, c->stream));
        P.cells = nullptr;
        P.cells_bar0 = b0;
        if (cell_table) {
            SMG_HIP(c, launch_cell_table(P.geom, b0, nb, P.sub, c->cell_tab_pts.ptr, c->stream));
            P.cells = c->cell_tab_pts.ptr;
        }
        SMG_HIP(c, launch_pe_points(P, pe_exact(c, prec), b0, nb, c->stream));
        A.n_pts = nb * npb;
        A.pt_bar0 = b0;
        A.accumulate = b0 > 0;
        SMG_HIP(c, scatter_row_cuts(c, A));
        SMG_HIP(c, launch_pair_terms(A, c->stream));
        SMG_HIP(c, scatter_partials(c, A));
        SMG_HIP(c, launch_scatter(A, prec == SMG_PREC_FP64, c->stream));
    }
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!dst_is_device)
        SMG_HIP(c, hipMemcpyAsync(dst, out, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)(g.n_det - 1) * (uint64_t)npb * (uint64_t)ncell;
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    return SMG_OK;
}

smg_status smg_pe_tangent(smg_ctx* c, int rotation, smg_precision prec, const double* dmu, float* dA,
                          float* A_out, int dst_is_device, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!dmu || !dA) return fail(c, SMG_ERR_BAD_PARAM, "dmu and dA must not be NULL");
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    const int n_par = g.n_layers + 2 * g.n_det;
    PeLaunch A{};
    A.geom = device_geom(c);
    A.dst = nullptr;
    A.counters = c->counters.ptr;
    const double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    pick_tile(g, A.tile, A.tiles);
    if ((s = ensure_spill(c, A)) != SMG_OK) return s;
    SMG_HIP(c, pe_cell_table(c, A));
    SMG_HIP(c, c->tan_v.reserve((size_t)n_par));
    SMG_HIP(c, hipMemcpyAsync(c->tan_v.ptr, dmu, n_par * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PeTangent TG{c->tan_v.ptr, dA, A_out};
    if (!dst_is_device) {
        SMG_HIP(c, c->out.reserve(n_el));
        TG.d_out = c->out.ptr;
        if (A_out) {
            SMG_HIP(c, c->out2.reserve(n_el));
            TG.a_out = c->out2.ptr;
        }
    }
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    SMG_HIP(c, launch_pe_tangent(A, TG, pe_exact(c, prec), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (!dst_is_device) {
        SMG_HIP(c, hipMemcpyAsync(dA, TG.d_out, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (A_out)
            SMG_HIP(c, hipMemcpyAsync(A_out, TG.a_out, n_el * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)(c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2]);
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    return SMG_OK;
}

smg_status smg_pe_backward(smg_ctx* c, int rotation, smg_precision prec, const float* w,
                           int w_is_device, double* grad, double* loss, smg_stats* stats) {
    double t_start = now_ms();
    smg_status s = check_ready(c, rotation);
    if (s != SMG_OK) return s;
    if (!w || !grad) return fail(c, SMG_ERR_BAD_PARAM, "w and grad must not be NULL");
    const GeomDesc& g = c->hg.g;
    const size_t n_el = (size_t)g.n_det * (size_t)g.n_vox;
    PeLaunch A{};
    A.geom = device_geom(c);
    A.dst = nullptr;
    A.counters = c->counters.ptr;
    double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    pick_tile(g, A.tile, A.tiles);
    if ((s = ensure_spill(c, A)) != SMG_OK) return s;
    SMG_HIP(c, pe_cell_table(c, A));
    const int n_tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
    PeGradScratch S;
    // enough waves to fill the GPU: about 8192 (bar, group of work items); a
    // tile's sub-cells are split into ranges (at least 8 sub-cells each) when
    // bars x tiles alone are fewer
    const int n_cells = A.sub[0] * A.sub[1] * A.sub[2];
    const int64_t bt = (int64_t)g.n_det * n_tiles;
    S.cell_slices = (int)std::max<int64_t>(1, std::min<int64_t>(n_cells / 8, (8192 + bt - 1) / bt));
    S.groups = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)n_tiles * S.cell_slices, (8192 + g.n_det - 1) / g.n_det));
    S.stride = (int64_t)g.n_det + g.n_layers + 2;
    SMG_HIP(c, c->scratch.reserve((size_t)g.n_det * S.groups * S.stride));
    S.rows = c->scratch.ptr;
    const int n_out = g.n_layers + 2 * g.n_det + 1;
    SMG_HIP(c, c->grad.reserve((size_t)n_out));
    if (w_is_device) {
        S.w = w;
    } else {
        SMG_HIP(c, c->in.reserve(n_el));
        SMG_HIP(c, hipMemcpyAsync(c->in.ptr, w, n_el * sizeof(float), hipMemcpyHostToDevice, c->stream));
        S.w = c->in.ptr;
    }
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    SMG_HIP(c, launch_pe_backward(A, S, pe_exact(c, prec), c->grad.ptr, c->stream));
    SMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    std::vector<double> out((size_t)n_out);
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SMG_HIP(c, hipMemcpyAsync(out.data(), c->grad.ptr, n_out * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipMemcpyAsync(cnt, c->counters.ptr, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    SMG_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    std::memcpy(grad, out.data(), (size_t)(n_out - 1) * sizeof(double));
    if (loss) *loss = out[(size_t)n_out - 1];
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = now_ms() - t_start;
        stats->elements = n_el;
        stats->rays = n_el * (uint64_t)(c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2]);
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    return SMG_OK;
}

}  // extern "C"
