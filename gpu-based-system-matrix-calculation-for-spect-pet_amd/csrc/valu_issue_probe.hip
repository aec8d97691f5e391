// Issue rate of single VALU instruction kinds on this GPU (gfx950): each kernel
// runs kChains independent dependency chains of one operation per lane, in
// registers only, and stores one value per thread at the end (out holds one
// element per thread of the grid).  Prints wave instructions per second and
// per SIMD per 4 cycles at 2.4 GHz (256 CUs x 4 SIMDs), the yardstick of the
// kernels' issue figures: 1.0 is one wave64 instruction every 4 cycles, the
// unpacked FMA rate.
//
//   bin/valu_issue_probe [blocks of 256 threads, default 4096]
//
// The instruction count per chain step of each kernel (ops column) was read
// from the compiler's assembly of this file.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 8;
constexpr int kIters = 4096;

#define PROBE_KERNEL(name, T, init, step)                                        \
    __global__ __launch_bounds__(256) void name(T a, T b, T* out) {               \
        T x[kChains];                                                             \
        _Pragma("unroll") for (int c = 0; c < kChains; ++c) x[c] = init;         \
        for (int i = 0; i < kIters; ++i) {                                        \
            _Pragma("unroll") for (int c = 0; c < kChains; ++c) { step; }        \
        }                                                                         \
        T s = x[0];                                                               \
        _Pragma("unroll") for (int c = 1; c < kChains; ++c) s += x[c];           \
        out[blockIdx.x * blockDim.x + threadIdx.x] = s;                           \
    }

// v_add_f32, v_mul_f32, v_fma_f32
PROBE_KERNEL(k_add_f32, float, a + (float)(threadIdx.x + c), x[c] = x[c] + b)
PROBE_KERNEL(k_mul_f32, float, a + (float)(threadIdx.x + c), x[c] = x[c] * b)
PROBE_KERNEL(k_fma_f32, float, a + (float)(threadIdx.x + c), x[c] = __builtin_fmaf(x[c], b, a))
// v_max_f32 + v_subrev_f32
PROBE_KERNEL(k_max_f32, float, a + (float)(threadIdx.x + c), x[c] = __builtin_fmaxf(x[c], b) - a)
// v_cmp_lt_f32 + v_cndmask_b32 + v_add_f32
PROBE_KERNEL(k_sel_f32, float, a + (float)(threadIdx.x + c), x[c] = x[c] > b ? x[c] - a : x[c] + a)
// v_rcp_f32 + v_add_f32
PROBE_KERNEL(k_rcp_f32, float, a + (float)(threadIdx.x + c), x[c] = __builtin_amdgcn_rcpf(x[c]) + b)
// v_cvt_f64_f32 + v_mul_f64 + v_cvt_f32_f64
PROBE_KERNEL(k_cvt_f32, float, a + (float)(threadIdx.x + c), x[c] = (float)((double)x[c] * 0.999999999999))
// v_add_f64, v_fma_f64
PROBE_KERNEL(k_add_f64, double, a + (double)(threadIdx.x + c), x[c] = x[c] + b)
PROBE_KERNEL(k_fma_f64, double, a + (double)(threadIdx.x + c), x[c] = __builtin_fma(x[c], b, a))
// v_xad_u32
PROBE_KERNEL(k_xad_u32, unsigned, (unsigned)a + threadIdx.x + c, x[c] = (x[c] ^ b) + 1u)

template <class T>
float time_ms(void (*k)(T, T, T*), T a, T b, T* out, int blocks) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0f;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, a, b, out);  // warm-up
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, a, b, out);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return -1.0f;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / 5.0f;
}

}  // namespace

int main(int argc, char** argv) {
    const int blocks = argc > 1 ? std::atoi(argv[1]) : 4096;
    if (blocks < 1 || blocks > (1 << 20)) {
        std::fprintf(stderr, "blocks must lie in [1, 2^20]\n");
        return 2;
    }
    void* out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)) != hipSuccess) return 1;
    const double waves = blocks * 4.0;
    const double peak = 256.0 * 4 * 2.4e9 / 4;  // wave instructions per second at 1 per 4 cycles
    std::printf("%-8s %5s %10s %14s %12s\n", "kernel", "ops", "ms", "wave-insts/s", "per 4 cyc");
    auto report = [&](const char* name, int ops, float ms) {
        if (ms <= 0.0f) {
            std::printf("%-8s failed\n", name);
            return;
        }
        const double wi = waves * kIters * kChains * ops;
        std::printf("%-8s %5d %10.3f %14.4e %12.3f\n", name, ops, ms, wi / (ms * 1e-3),
                    wi / (ms * 1e-3) / peak);
        std::fflush(stdout);
    };
    float* of = static_cast<float*>(out);
    double* od = static_cast<double*>(out);
    unsigned* ou = static_cast<unsigned*>(out);
    report("add_f32", 1, time_ms(k_add_f32, 1.0f, 1.0000001f, of, blocks));
    report("mul_f32", 1, time_ms(k_mul_f32, 1.0f, 0.9999999f, of, blocks));
    report("fma_f32", 1, time_ms(k_fma_f32, 1e-7f, 0.9999999f, of, blocks));
    report("max_f32", 2, time_ms(k_max_f32, 1e-7f, 0.5f, of, blocks));
    report("sel_f32", 3, time_ms(k_sel_f32, 1e-7f, 0.5f, of, blocks));
    report("rcp_f32", 2, time_ms(k_rcp_f32, 1.5f, 0.5f, of, blocks));
    report("cvt_f32", 3, time_ms(k_cvt_f32, 1.0f, 0.9999999f, of, blocks));
    report("add_f64", 1, time_ms(k_add_f64, 1.0, 1.0000000001, od, blocks));
    report("fma_f64", 1, time_ms(k_fma_f64, 1e-9, 0.999999999, od, blocks));
    report("xad_u32", 1, time_ms(k_xad_u32, 1u, 0x5555u, ou, blocks));
    (void)hipFree(out);
    return 0;
}
