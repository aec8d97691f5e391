// Placeholders for entry points still being written (removed once implemented).
#include <hip/hip_runtime.h>
#include "../../include/smg.h"
extern "C" {
smg_status smg_pe_backward(smg_ctx*, int, const float*, int, double*, double*, smg_stats*) { return SMG_ERR_STATE; }
int64_t smg_bitmap_crystal_words(int n) { int64_t b = (int64_t)n * (n + 1) / 2 * n; return (b + 31) / 32; }
int64_t smg_bitmap_collimator_words(int h, int n) { int64_t b = (int64_t)h * n * n; return (b + 31) / 32; }
smg_status smg_bitmap_crystal(smg_ctx*, uint32_t*, int64_t, int) { return SMG_ERR_STATE; }
smg_status smg_bitmap_collimator(smg_ctx*, uint32_t*, int64_t, int) { return SMG_ERR_STATE; }
smg_status smg_load_bitmap_crystal(smg_ctx*, const uint32_t*, int64_t) { return SMG_ERR_STATE; }
smg_status smg_load_bitmap_collimator(smg_ctx*, const uint32_t*, int64_t) { return SMG_ERR_STATE; }
smg_status smg_scatter_crystal(smg_ctx*, int, smg_precision, const float*, int, float*, int, smg_stats*) { return SMG_ERR_STATE; }
smg_status smg_scatter_collimator(smg_ctx*, int, smg_coll_variant, smg_precision, float*, int, smg_stats*) { return SMG_ERR_STATE; }
}
