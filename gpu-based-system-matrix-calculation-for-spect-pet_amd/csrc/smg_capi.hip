// C ABI of the system-matrix generator: context, geometry upload, launches.
//
// Replaces the per-call device setup of the reference launch wrappers
// (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1368-1437,
// ScatterGen_RayTracing_CircularHole/scatter.cu:2013-2378): a context keeps the
// parsed geometry, the bitmaps and staging buffers on its GPU across calls
// (rotations, calibration steps), and reports errors as status codes instead of
// exiting the process.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
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

struct smg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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
    DevBuf<float> out;        // staging for host outputs
    DevBuf<float> in;         // staging for host inputs (PE weights)
    DevBuf<unsigned long long> counters;
    DevBuf<uint32_t> bm_crystal, bm_coll;
    bool have_bm_crystal = false, have_bm_coll = false;
    DevBuf<double> scratch;   // adjoint partials
    DevBuf<double> grad;
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

smg_status check_ready(smg_ctx* c, int rotation) {
    if (!c) return SMG_ERR_BAD_PARAM;
    if (!c->have_geom) return fail(c, SMG_ERR_STATE, "no geometry set");
    if (rotation < 0) return fail(c, SMG_ERR_BAD_PARAM, "negative rotation index %d", rotation);
    SMG_HIP(c, hipSetDevice(c->device));
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        c->counters.reserve(4) != hipSuccess) {
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
    c->out.release();
    c->in.release();
    c->counters.release();
    c->bm_crystal.release();
    c->bm_coll.release();
    c->scratch.release();
    c->grad.release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
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
    if ((s = upload(c, c->bars_f, hg.bars_f)) != SMG_OK) return s;
    if ((s = upload(c, c->bars_d, hg.bars_d)) != SMG_OK) return s;
    if ((s = upload(c, c->holes_f, hg.holes_f)) != SMG_OK) return s;
    if ((s = upload(c, c->holes_d, hg.holes_d)) != SMG_OK) return s;
    if ((s = upload(c, c->layers_f, hg.layers_f)) != SMG_OK) return s;
    if ((s = upload(c, c->layers_d, hg.layers_d)) != SMG_OK) return s;
    SMG_HIP(c, hipStreamSynchronize(c->stream));
    c->hg = std::move(hg);
    c->have_geom = true;
    c->have_bm_crystal = c->have_bm_coll = false;
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
    PeLaunch A;
    A.geom = device_geom(c);
    A.dst = out;
    A.counters = c->counters.ptr;
    double phi = (double)rotation * g.dphi;
    A.cphi = std::cos(phi);
    A.sphi = std::sin(phi);
    for (int k = 0; k < 3; ++k) A.sub[k] = c->pe_sub[k];
    pick_tile(g, A.tile, A.tiles);
    SMG_HIP(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), c->stream));
    SMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    SMG_HIP(c, launch_pe_forward(A, prec == SMG_PREC_FP64, c->stream));
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
        stats->rays = n_el * (uint64_t)(c->pe_sub[0] * c->pe_sub[1] * c->pe_sub[2]);
        stats->fp64_fallbacks = cnt[0];
        stats->overflow_waves = cnt[1];
    }
    if (cnt[1])
        return fail(c, SMG_ERR_CAPACITY,
                    "%llu waves exceeded the candidate capacity (%d bars, %d holes)", cnt[1],
                    kCapBars, kCapHoles);
    return SMG_OK;
}

}  // extern "C"
