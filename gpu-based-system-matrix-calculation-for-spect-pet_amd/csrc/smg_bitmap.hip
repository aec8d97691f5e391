// Geometry bitmaps of the scatter kernels, gfx950.
//
// Crystal bitmap (ScatterGen_RayTracing_CircularHole/scatter.cu:1440-1540):
// bit (i<j, k) says bar k lies between bars i and j by an angular
// bounding-sphere test seen from i or from j; bits of the unordered pair
// {i, j} start at (i N - i(i+1)/2 + j) N (scatter.cu:1040-1059).
// Collimator bitmap (scatter.cu:1939-2009): bit (h, j, k) says bar k lies on
// the ray from layer-0 hole h toward bar j; bit index h N^2 + j N + k.
//
// The reference runs one thread per bit and sets bits with atomicOr.  Here one
// thread owns one 32-bit word: it evaluates its 32 predicates and stores the
// word once (no atomics, coalesced stores).  Predicates run in FP32 and are
// re-evaluated in FP64 when within rounding distance of their threshold, so
// the words match an FP64 evaluation of the reference test.
#include <hip/hip_runtime.h>

#include "smg_kernels.hpp"

namespace smg {

namespace {

template <class T>
__device__ __forceinline__ T dist3(T ax, T ay, T az, T bx, T by, T bz) {
    T dx = ax - bx, dy = ay - by, dz = az - bz;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// |u_{k<-i} - u_{j<-i}| <= R_k / L_ik + R_j / L_ij, the test seen from i.
template <class T>
__device__ __forceinline__ bool seen_from(const Sphere<T>& i, const Sphere<T>& j,
                                          const Sphere<T>& k, bool& near) {
    T lij = dist3(i.x, i.y, i.z, j.x, j.y, j.z);
    T lik = dist3(k.x, k.y, k.z, i.x, i.y, i.z);
    T ux = (k.x - i.x) / lik - (j.x - i.x) / lij;
    T uy = (k.y - i.y) / lik - (j.y - i.y) / lij;
    T uz = (k.z - i.z) / lik - (j.z - i.z) / lij;
    T s = sqrt(ux * ux + uy * uy + uz * uz);
    T lim = k.r / lik + j.r / lij;
    return g_le(s, lim, T(64) * kUnit<T> * (s + lim + T(1)), near);
}

// Crystal predicate: from i, else from j (scatter.cu:1518-1533).
template <class T>
__device__ __forceinline__ bool crystal_pred(const Sphere<T>& i, const Sphere<T>& j,
                                             const Sphere<T>& k, bool& near) {
    bool n1 = false;
    bool a = seen_from(i, j, k, n1);
    near |= n1;
    if (a && !n1) return true;
    bool n2 = false;
    bool b = seen_from(j, i, k, n2);
    near |= n2;
    return a || b;
}

__device__ __forceinline__ bool crystal_bit(const BitmapLaunch& A, int i, int j, int k) {
    bool near = false;
    bool r = crystal_pred<float>(A.spheres_f[i], A.spheres_f[j], A.spheres_f[k], near);
    if (!near) return r;
    bool n2 = false;
    return crystal_pred<double>(A.spheres_d[i], A.spheres_d[j], A.spheres_d[k], n2);
}

__device__ __forceinline__ bool coll_bit(const BitmapLaunch& A, int h, int j, int k) {
    bool near = false;
    bool r = seen_from<float>(A.holes_f[h], A.spheres_f[j], A.spheres_f[k], near);
    if (!near) return r;
    bool n2 = false;
    return seen_from<double>(A.holes_d[h], A.spheres_d[j], A.spheres_d[k], n2);
}

// First pair index of row i (the pair (i, i)): S(i) = i N - i(i-1)/2.
__device__ __forceinline__ int64_t row_start(int64_t i, int64_t n) { return i * n - i * (i - 1) / 2; }

__global__ __launch_bounds__(256) void bitmap_crystal_kernel(BitmapLaunch A) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= A.n_words) return;
    const int64_t n = A.n_det;
    const int64_t total = n * (n + 1) / 2 * n;
    int64_t b = w * 32;
    int64_t pair = b / n;
    int k = (int)(b % n);
    // invert pair = S(i) + (j - i)
    double disc = (2.0 * n + 1) * (2.0 * n + 1) - 8.0 * (double)pair;
    int64_t i = (int64_t)floor(((2.0 * n + 1) - sqrt(disc > 0 ? disc : 0.0)) / 2.0);
    if (i < 0) i = 0;
    if (i > n - 1) i = n - 1;
    while (i > 0 && row_start(i, n) > pair) --i;
    while (i + 1 < n && row_start(i + 1, n) <= pair) ++i;
    int64_t j = pair - row_start(i, n) + i;
    uint32_t word = 0;
    for (int bit = 0; bit < 32 && b < total; ++bit, ++b) {
        if (i < j && k != i && k != j && crystal_bit(A, (int)i, (int)j, k)) word |= 1u << bit;
        if (++k == n) {
            k = 0;
            if (++j == n) {
                ++i;
                j = i;
            }
        }
    }
    A.words[w] = word;
}

__global__ __launch_bounds__(256) void bitmap_coll_kernel(BitmapLaunch A) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= A.n_words) return;
    const int64_t n = A.n_det;
    const int64_t total = (int64_t)A.n_holes * n * n;
    int64_t b = w * 32;
    int h = (int)(b / (n * n));
    int64_t rem = b % (n * n);
    int j = (int)(rem / n), k = (int)(rem % n);
    uint32_t word = 0;
    for (int bit = 0; bit < 32 && b < total; ++bit, ++b) {
        if (coll_bit(A, h, j, k)) word |= 1u << bit;
        if (++k == n) {
            k = 0;
            if (++j == n) {
                j = 0;
                ++h;
            }
        }
    }
    A.words[w] = word;
}

}  // namespace

hipError_t launch_bitmap(const BitmapLaunch& A, bool collimator, hipStream_t stream) {
    if (A.n_words <= 0) return hipSuccess;
    const int64_t blocks = (A.n_words + 255) / 256;
    if (collimator)
        hipLaunchKernelGGL(bitmap_coll_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, A);
    else
        hipLaunchKernelGGL(bitmap_crystal_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, A);
    return hipGetLastError();
}

}  // namespace smg
