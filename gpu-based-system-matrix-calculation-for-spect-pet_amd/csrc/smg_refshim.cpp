// Reference-signature entry points (PESysMatGen, the three scatter()
// overloads) on top of the C ABI.  See include/smg_reference.h.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/smg.h"
#include "../../include/smg_reference.h"

namespace {

[[noreturn]] void die(smg_ctx* ctx, smg_status s, const char* where) {
    std::fprintf(stderr, "%s failed: %s: %s\n", where, smg_status_string(s),
                 ctx ? smg_last_error(ctx) : "");
    std::exit(EXIT_FAILURE);
}

void check(smg_ctx* ctx, smg_status s, const char* where) {
    if (s != SMG_OK) die(ctx, s, where);
}

void parse_triple(const char* env, int def[3]) {
    const char* v = std::getenv(env);
    if (!v) return;
    int a, b, c;
    if (std::sscanf(v, "%d,%d,%d", &a, &b, &c) == 3 && a > 0 && b > 0 && c > 0) {
        def[0] = a;
        def[1] = b;
        def[2] = c;
    } else {
        std::fprintf(stderr, "ignoring %s=%s (expected sx,sy,sz)\n", env, v);
    }
}

smg_precision precision() {
    const char* v = std::getenv("SMG_PRECISION");
    return (v && std::strcmp(v, "fp64") == 0) ? SMG_PREC_FP64 : SMG_PREC_FP32;
}

// One context per device, kept for the life of the process; the geometry is
// re-uploaded only when the parameter arrays change (img[20], the rotation
// index, does not count).
struct Cached {
    smg_ctx* ctx = nullptr;
    std::vector<float> coll, det, img, phys;
};

std::map<int, Cached>& cache() {
    static std::map<int, Cached> m;
    return m;
}

int pick_device(int cuda_id) {
    int n = 0;
    smg_device_count(&n);
    if (n <= 0) {
        std::fprintf(stderr, "no GPU available\n");
        std::exit(EXIT_FAILURE);
    }
    if (cuda_id < 0 || cuda_id >= n) {
        std::printf("cuda_id > the number of GPUs on the host! Set Device to Device 0!\n");
        return 0;
    }
    return cuda_id;
}

smg_ctx* context_for(int cuda_id, const float* coll, const float* det, const float* img,
                     const float* phys) {
    static const std::vector<float> zeros_coll(80000, 0.0f), zeros_phys(100, 0.0f);
    if (!coll) coll = zeros_coll.data();
    if (!phys) phys = zeros_phys.data();
    int dev = pick_device(cuda_id);
    Cached& c = cache()[dev];
    if (!c.ctx) check(nullptr, smg_create(dev, &c.ctx), "smg_create");
    std::vector<float> img_key(img, img + 100);
    img_key[20] = 0.0f;
    bool same = !c.coll.empty() && std::memcmp(c.coll.data(), coll, 80000 * sizeof(float)) == 0 &&
                std::memcmp(c.det.data(), det, 80000 * sizeof(float)) == 0 &&
                c.img == img_key && std::memcmp(c.phys.data(), phys, 100 * sizeof(float)) == 0;
    if (!same) {
        check(c.ctx, smg_set_geometry(c.ctx, coll, 80000, det, 80000, img, 100, phys, 100),
              "smg_set_geometry");
        c.coll.assign(coll, coll + 80000);
        c.det.assign(det, det + 80000);
        c.img = img_key;
        c.phys.assign(phys, phys + 100);
    }
    int pe[3] = {8, 16, 8}, sc[3] = {1, 1, 1};
    parse_triple("SMG_PE_SUBSAMPLES", pe);
    parse_triple("SMG_SCATTER_SUBSAMPLES", sc);
    check(c.ctx, smg_set_pe_subsamples(c.ctx, pe[0], pe[1], pe[2]), "smg_set_pe_subsamples");
    check(c.ctx, smg_set_scatter_subsamples(c.ctx, sc[0], sc[1], sc[2]), "smg_set_scatter_subsamples");
    return c.ctx;
}

// img[20] is the rotation index the drivers write (PEGen_CircularHole.cpp:133);
// the reference multiplies it, as a float, by the angle step
// (PESysMatGen.cu:1004-1006).  The C ABI takes whole indices, so an index
// that is not a non-negative whole number is refused rather than rounded.
int rotation_of(const float* img) {
    const float r = img[20];
    if (!(r >= 0.0f) || r != std::floor(r) || r > 1.0e6f) {
        std::fprintf(stderr, "rotation index img[20] = %g is not a non-negative whole number\n", (double)r);
        std::exit(EXIT_FAILURE);
    }
    return (int)r;
}

int64_t n_voxels(smg_ctx* ctx, int* n_det = nullptr, int* n_holes = nullptr) {
    int nd = 0, nh = 0, nr = 0;
    int64_t nv = 0;
    check(ctx, smg_get_dims(ctx, &nd, &nv, &nr, &nh), "smg_get_dims");
    if (n_det) *n_det = nd;
    if (n_holes) *n_holes = nh;
    return nv;
}

// Compute the bitmap and write the cache file (flag == 1), or read it.
void bitmap_from_flag(smg_ctx* ctx, bool collimator, int flag, const char* fname) {
    int nd = 0, nh = 0;
    n_voxels(ctx, &nd, &nh);
    int64_t words = collimator ? smg_bitmap_collimator_words(nh, nd) : smg_bitmap_crystal_words(nd);
    std::vector<uint32_t> buf((size_t)(words > 0 ? words : 1), 0u);
    if (flag == 1) {
        check(ctx, collimator ? smg_bitmap_collimator(ctx, buf.data(), words, 0)
                              : smg_bitmap_crystal(ctx, buf.data(), words, 0),
              "geometry bitmap");
        FILE* f = std::fopen(fname, "wb+");
        if (!f || std::fwrite(buf.data(), sizeof(uint32_t), (size_t)words, f) != (size_t)words) {
            std::fprintf(stderr, "cannot write geometry bitmap %s\n", fname);
            std::exit(EXIT_FAILURE);
        }
        std::fclose(f);
    } else {
        FILE* f = std::fopen(fname, "rb");
        if (!f) {
            std::fprintf(stderr, "cannot open geometry bitmap %s\n", fname);
            std::exit(EXIT_FAILURE);
        }
        size_t got = std::fread(buf.data(), sizeof(uint32_t), (size_t)words, f);
        std::fclose(f);
        if (got != (size_t)words) {
            std::fprintf(stderr, "geometry bitmap %s holds %zu words, %lld expected\n", fname, got,
                         (long long)words);
            std::exit(EXIT_FAILURE);
        }
        check(ctx, collimator ? smg_load_bitmap_collimator(ctx, buf.data(), words)
                              : smg_load_bitmap_crystal(ctx, buf.data(), words),
              "load geometry bitmap");
    }
}

}  // namespace

int PESysMatGen(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
                float* dst, int cuda_id) {
    smg_ctx* ctx = context_for(cuda_id, parameter_Collimator, parameter_Detector, parameter_Image,
                               nullptr);
    int64_t nv = n_voxels(ctx);
    smg_stats st;
    check(ctx, smg_pe_forward(ctx, rotation_of(parameter_Image), precision(), dst, 0, &st),
          "smg_pe_forward");
    std::printf("PE kernel: %.1f ms, %llu elements, %llu FP64 fallback rays\n", st.kernel_ms,
                (unsigned long long)st.elements, (unsigned long long)st.fp64_fallbacks);
    return (int)nv;
}

// Extension behind the three scatter() overloads (declared in smg_reference.h):
// the scatter block of `which` and, in the same call on the GPU, PE + scatter.
int scatter_and_sum(int which, float* parameter_Collimator, float* parameter_Detector,
                    float* parameter_Image, float* parameter_Physics, float* PE_SysMat, float* PE_add,
                    const char* FnameGeoCrystal, const char* FnameGeoCollimator, float* scatter_dst,
                    float* sum_dst, int cuda_id) {
    if (which < 1 || which > 3 || (!scatter_dst && !sum_dst)) {
        std::fprintf(stderr, "scatter_and_sum: which = %d, outputs %p %p\n", which, (void*)scatter_dst,
                     (void*)sum_dst);
        std::exit(EXIT_FAILURE);
    }
    smg_ctx* ctx = context_for(cuda_id, which == 1 ? nullptr : parameter_Collimator, parameter_Detector,
                               parameter_Image, parameter_Physics);
    int64_t nv = n_voxels(ctx);
    unsigned parts = 0;
    if (which & 1) {
        bitmap_from_flag(ctx, false, (int)std::floor(parameter_Physics[8]), FnameGeoCrystal);
        parts |= SMG_PART_CRYSTAL;
    }
    if (which & 2) {
        bitmap_from_flag(ctx, true, (int)std::floor(parameter_Physics[9] + 0.001f), FnameGeoCollimator);
        parts |= which == 3 ? SMG_PART_COLL_COMBINED : SMG_PART_COLL_STANDALONE;
    }
    smg_stats st;
    check(ctx, smg_scatter_sum(ctx, rotation_of(parameter_Image), parts, precision(), PE_SysMat, PE_add, 0,
                               scatter_dst, sum_dst, 0, &st),
          "smg_scatter_sum");
    std::printf("%s scatter%s: %.1f ms\n", which == 3 ? "crystal + collimator" : (which == 1 ? "crystal" : "collimator"),
                sum_dst ? " and PE + scatter" : "", st.kernel_ms);
    return (int)nv;
}

// Combined generator: crystal scatter + collimator scatter (r = 0.1 mm), summed
// on the GPU.
int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
            float* parameter_Physics, float* PE_SysMat, const char* FnameGeoCrystal,
            const char* FnameGeoCollimator, float* dst, int cuda_id) {
    return scatter_and_sum(3, parameter_Collimator, parameter_Detector, parameter_Image, parameter_Physics,
                           PE_SysMat, nullptr, FnameGeoCrystal, FnameGeoCollimator, dst, nullptr, cuda_id);
}

// Standalone collimator scatter (r = 0.2 mm virtual disc).
int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image,
            float* parameter_Physics, float* PE_SysMat, const char* FnameGeo, float* dst,
            int cuda_id) {
    (void)PE_SysMat;  // loaded by the reference driver but not used by its kernel
    smg_ctx* ctx = context_for(cuda_id, parameter_Collimator, parameter_Detector, parameter_Image,
                               parameter_Physics);
    int64_t nv = n_voxels(ctx);
    bitmap_from_flag(ctx, true, (int)std::floor(parameter_Physics[9] + 0.001f), FnameGeo);
    smg_stats st;
    check(ctx, smg_scatter_collimator(ctx, rotation_of(parameter_Image), SMG_COLL_STANDALONE,
                                      precision(), dst, 0, &st),
          "smg_scatter_collimator");
    std::printf("collimator scatter: %.1f ms\n", st.kernel_ms);
    return (int)nv;
}

// Standalone crystal scatter.
int scatter(float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeo, float* dst, int cuda_id) {
    smg_ctx* ctx = context_for(cuda_id, nullptr, parameter_Detector, parameter_Image,
                               parameter_Physics);
    int64_t nv = n_voxels(ctx);
    bitmap_from_flag(ctx, false, (int)std::floor(parameter_Physics[8]), FnameGeo);
    smg_stats st;
    check(ctx, smg_scatter_crystal(ctx, rotation_of(parameter_Image), precision(), PE_SysMat, 0,
                                   dst, 0, &st),
          "smg_scatter_crystal");
    std::printf("crystal scatter: %.1f ms\n", st.kernel_ms);
    return (int)nv;
}
