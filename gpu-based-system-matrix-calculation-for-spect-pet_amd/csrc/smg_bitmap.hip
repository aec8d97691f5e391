// Geometry bitmaps of the scatter kernels, gfx950.
//
// The reference runs one thread per bit and sets bits with atomicOr
// (scatter.cu:1538).  Here one wave owns 64 consecutive 32-bit words (2,048
// bits): at step s lane l evaluates bit 64 s + l of the wave's run, so the
// lanes read consecutive bars k, and a ballot gives words 2s and 2s + 1, which
// lanes 2s and 2s + 1 keep; after 32 steps each lane stores its word (no
// atomics, one coalesced 256-byte store per wave).  A lane keeps its row's
// terms across steps and recomputes them only when its k wraps into the next
// row.  The bit predicates and the row walk are in smg_bitmap_body.hpp.
// A grid-stride loop over the runs covers any bitmap size (the reference's
// limit of N = 6,665 bars gives 4.6e9 words).
#include <hip/hip_runtime.h>

#include "smg_bitmap_body.hpp"
#include "smg_kernels.hpp"

namespace smg {

namespace {

constexpr int kWordsPerRun = 64;                // one word per lane
constexpr int64_t kBitsPerRun = 32 * kWordsPerRun;

template <bool kColl>
__global__ __launch_bounds__(256) void bitmap_kernel(BitmapLaunch A) {
    const int lane = threadIdx.x & 63;
    const int64_t n = A.n_det;
    const int64_t total = bm_total_bits(kColl, n, A.n_holes);
    const int64_t n_runs = (A.n_words + kWordsPerRun - 1) / kWordsPerRun;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    // the run index is the same for the 64 lanes of a wave, so every ballot
    // below sees the whole wave
    for (int64_t run = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6; run < n_runs;
         run += stride) {
        int64_t b = run * kBitsPerRun + lane;
        BitRow r{};
        int k = 0;
        if (b < total) {
            bm_row_of_bit(kColl, b, n, r, k);
            bm_row_setup(kColl, A.spheres_d, A.holes_d, r);
        }
        uint32_t word = 0;
        for (int s = 0; s < 32; ++s) {
            const bool bit = b < total && bm_bit(kColl, r, A.spheres_d, A.holes_d, k);
            const uint64_t m = __ballot(bit);
            if (lane == 2 * s) word = (uint32_t)m;
            if (lane == 2 * s + 1) word = (uint32_t)(m >> 32);
            b += 64;
            k += 64;
            if (b < total && k >= n) {
                do {
                    k -= (int)n;
                    bm_next_row(kColl, n, r);
                } while (k >= n);
                bm_row_setup(kColl, A.spheres_d, A.holes_d, r);
            }
        }
        const int64_t w = run * kWordsPerRun + lane;
        if (w < A.n_words) A.words[w] = word;
    }
}

}  // namespace

hipError_t launch_bitmap(const BitmapLaunch& A, bool collimator, hipStream_t stream) {
    if (A.n_words <= 0) return hipSuccess;
    const int64_t n_runs = (A.n_words + kWordsPerRun - 1) / kWordsPerRun;
    // four runs per block of 256 threads; at most 2^16 blocks (262,144 waves),
    // the rest by the grid-stride loop
    int64_t blocks = (n_runs + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    if (collimator)
        hipLaunchKernelGGL(bitmap_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, stream, A);
    else
        hipLaunchKernelGGL(bitmap_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, stream, A);
    return hipGetLastError();
}

}  // namespace smg
