/*
 * smg: C ABI of the MI355X (gfx950) SPECT/PET system-matrix generator.
 *
 * Plain C types only: float/double/uint32 pointers and sizes.  One context is
 * bound to one GPU and one HIP stream; calls on a context are synchronous and
 * must come from one host thread at a time.  A pointer argument flagged
 * "_is_device" must be device memory of the context's GPU (for example a
 * hipMalloc'ed buffer or a torch tensor's data_ptr()); otherwise it is host
 * memory and the call copies through the context's own device buffers.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   PESysMatGen()  PEGen_RayTracing_CircularHole/PESysMatGen.h:4-5,
 *                  defined PEGen_RayTracing_CircularHole/PESysMatGen.cu:1368-1437
 *   scatter()      ScatterGen_RayTracing_CircularHole/scatter.h:4-5 (combined),
 *                  ScatterGen_Collimator/scatter.h:4-5, ScatterGen_Crystal/scatter.h:4
 * The C++ shims with the reference's exact signatures are declared in
 * smg_reference.h and built on top of these entry points.
 *
 * Matrix layout (all outputs): FP32 [detector][voxel], voxel index
 * ix + iy*NX + iz*NX*NY, one rotation per call
 * (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1022-1034).
 */
#ifndef SMG_H
#define SMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMG_ABI_VERSION 1

typedef enum smg_status {
    SMG_OK = 0,
    SMG_ERR_BAD_PARAM = 1,  /* invalid argument or parameter array content */
    SMG_ERR_CAPACITY = 2,   /* geometry beyond a fixed capacity */
    SMG_ERR_IO = 3,         /* file could not be read or written */
    SMG_ERR_HIP = 4,        /* HIP runtime error */
    SMG_ERR_NO_DEVICE = 5,  /* no GPU, or device index out of range */
    SMG_ERR_STATE = 6       /* call order: e.g. no geometry set yet */
} smg_status;

typedef enum smg_precision {
    SMG_PREC_FP32 = 0,  /* FP32 arithmetic, compensated sums, FP64 for
                           near-threshold and near-grazing rays */
    SMG_PREC_FP64 = 1   /* all-FP64 arithmetic */
} smg_precision;

typedef enum smg_coll_variant {
    SMG_COLL_COMBINED = 0,   /* r = 0.1 mm virtual disc, cos factor
                                (ScatterGen_RayTracing_CircularHole/scatter.cu:1702-1716) */
    SMG_COLL_STANDALONE = 1  /* r = 0.2 mm, (r^2 + T r)/(4 L^2)
                                (ScatterGen_Collimator/scatter.cu:1207-1215) */
} smg_coll_variant;

/* Parts of smg_scatter_sum (bit mask). */
typedef enum smg_scatter_part {
    SMG_PART_CRYSTAL = 1,          /* crystal scatter (smg_scatter_crystal) */
    SMG_PART_COLL_COMBINED = 2,    /* collimator scatter, r = 0.1 mm (combined driver) */
    SMG_PART_COLL_STANDALONE = 4   /* collimator scatter, r = 0.2 mm (standalone driver) */
} smg_scatter_part;

typedef struct smg_stats {
    double kernel_ms;          /* GPU time of the main kernel(s), HIP events */
    double total_ms;           /* host wall time of the call */
    uint64_t elements;         /* matrix elements produced */
    uint64_t rays;             /* element x sub-cell (x scatterer) terms */
    uint64_t fp64_fallbacks;   /* FP32 terms recomputed in FP64 */
    uint64_t overflow_waves;   /* work items whose candidate lists overflowed LDS and were
                                  redone with lists in global memory (same result) */
} smg_stats;

typedef struct smg_ctx smg_ctx;

/* ---- context --------------------------------------------------------- */
smg_status smg_device_count(int* count);
smg_status smg_create(int device, smg_ctx** ctx);
smg_status smg_destroy(smg_ctx* ctx);
const char* smg_last_error(const smg_ctx* ctx);
const char* smg_status_string(smg_status status);
/* The context's stream (a hipStream_t) for callers that time or order work. */
void* smg_stream(smg_ctx* ctx);
smg_status smg_synchronize(smg_ctx* ctx);

/* ---- geometry ----------------------------------------------------------
 * The four parameter arrays of Params_{Collimator,Detector,Image,Physics}.dat
 * (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:33-56): at least
 * 80000, 80000, 100 and 100 floats.  Copied into the context (FP32 and FP64
 * records on the device).  img[20] (rotation index) is ignored: the rotation
 * is an explicit argument of each entry point. */
smg_status smg_set_geometry(smg_ctx* ctx, const float* coll, size_t n_coll, const float* det,
                            size_t n_det, const float* img, size_t n_img, const float* phys,
                            size_t n_phys);
/* Sizes parsed from the geometry: bars N, voxels V, rotations R, layer-0 holes. */
smg_status smg_get_dims(const smg_ctx* ctx, int* n_det, int64_t* n_vox, int* n_rot,
                        int* n_holes_l0);
/* Sub-cells per bar of the PE matrix.  The reference hard-codes 8 x 16 x 8
 * (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1037), the default here. */
smg_status smg_set_pe_subsamples(smg_ctx* ctx, int sx, int sy, int sz);
/* Detection sub-cells of the scatter matrices; the reference uses 1 x 1 x 1
 * (ScatterGen_RayTracing_CircularHole/scatter.cu:1140), the default here. */
smg_status smg_set_scatter_subsamples(smg_ctx* ctx, int sx, int sy, int sz);
/* mu(E) tables for the Compton kernels' second leg and absorption at E'.  No
 * counterpart in the reference, which scales each coefficient from E0 by
 * (E0/E')^3 (photoelectric) and E0/E' (Compton) (ScatterGen_RayTracing_
 * CircularHole/scatter.cu:1371-1375).  With tables, each bar's coefficients
 * at E0 are scaled by mu(E')/mu(E0) of the crystal tables and the collimator
 * coefficients by the collimator tables; interpolation is linear in
 * (log E, log mu) between nodes and continues the end segments beyond them.
 * n_nodes: 2..256 increasing energies in keV, positive values; 0 restores the
 * power laws (the default). */
smg_status smg_set_mu_tables(smg_ctx* ctx, int n_nodes, const double* energy_kev,
                             const double* crystal_pe, const double* crystal_compton,
                             const double* coll_pe, const double* coll_compton);

/* Elliptical holes.  The reference defines the elliptic-cylinder chord
 * length_ellipticalcylinder_ray (PEGen_RayTracing_CircularHole/PESysMatGen.cu:
 * 801-947) but never calls it, and its parameter file has no fields for it;
 * this entry point supplies them.  shapes: 3 floats per hole of the current
 * geometry (n_holes = all layers' holes, in parameter order): x_rho, z_rho
 * (semi-axes, mm) and theta (radians, the turn of the ellipse's axes about
 * the hole axis, as the reference's routine takes it); a hole with x_rho <= 0
 * keeps the parameter file's circle.  shapes = NULL makes every hole a circle.
 * The PE matrix, its adjoint, the projections and the crystal grid weights
 * trace elliptical holes with the reference routine's chord: the matrix and
 * the projections in the requested precision (FP32 kernels with the elliptic
 * chord compiled in), the adjoint, tangent and crystal grid weights on their
 * all-FP64 kernels whatever the request; culling and the candidate scans use
 * the bounding radius max(x_rho, z_rho).  The collimator grid scatterers
 * (smg_scatter_collimator_grid) leave out the grid points inside an ellipse
 * and subtract the elliptic-cylinder chords from the plate material.
 * smg_set_geometry resets every hole to a circle. */
smg_status smg_set_hole_shapes(smg_ctx* ctx, const float* shapes, int n_holes);

/* ---- photoelectric matrix (replaces PESysMatGen) ---------------------------
 * Writes the N x V block of rotation `rotation` to dst. */
smg_status smg_pe_forward(smg_ctx* ctx, int rotation, smg_precision prec, float* dst,
                          int dst_is_device, smg_stats* stats);

/* Tangent (forward-mode derivative) of smg_pe_forward along a direction dmu
 * over the parameters of smg_pe_backward's gradient (dmu[0 .. Lc-1] layers,
 * dmu[Lc .. Lc+N-1] mu_total and dmu[Lc+N .. Lc+2N-1] mu_pe of each bar, host
 * memory): dA[d][v] = sum_k dmu[k] dA[d][v]/dmu_k, N x V floats, and the
 * matrix itself into A (optional, NULL to skip).  With smg_pe_backward it
 * gives Gauss-Newton products J^T (J v) at two passes per product, whatever
 * the number of parameters.  No counterpart in the reference. */
smg_status smg_pe_tangent(smg_ctx* ctx, int rotation, smg_precision prec, const double* dmu, float* dA,
                          float* A, int dst_is_device, smg_stats* stats);

/* smg_pe_forward with no allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  dst_device: a device
 * buffer of N x V floats (the caller's); the geometry and sub-cells must be
 * set beforehand (those calls allocate).  Same kernels and results as
 * smg_pe_forward.  Its counters and sub-cell table live in a fixed 32 MiB
 * scratch array per GPU, held until the process ends and shared by every
 * context on that GPU, so calls that overlap on one GPU race.  Limits
 * (SMG_ERR_CAPACITY): bars x sub-cells x 136 bytes must fit in 32 MiB - 256
 * bytes (2048 bars at 3 x 3 x 3 sub-cells take 7.5 MB), and a voxel tile
 * that needs more than 448 bars or 256 holes of candidates is refused (its
 * elements are then left unwritten; smg_pe_forward handles such tiles). */
smg_status smg_pe_forward_noalloc(smg_ctx* ctx, int rotation, smg_precision prec, float* dst_device,
                                  smg_stats* stats);

/* dL/dmu of L = sum_dv w[d][v] A[d][v] (adjoint of smg_pe_forward), for the
 * mu of each collimator layer, mu_total and mu_pe of each bar:
 * grad[0 .. Lc-1], grad[Lc .. Lc+N-1], grad[Lc+N .. Lc+2N-1].  w is N x V.
 * loss (optional) receives L.  prec selects the ray arithmetic as in
 * smg_pe_forward (FP32 with FP64 re-computation near thresholds, or all
 * FP64); the sums are FP64 either way.  Deterministic: fixed-order
 * reductions.  No counterpart in the reference, which has no gradient. */
smg_status smg_pe_backward(smg_ctx* ctx, int rotation, smg_precision prec, const float* w,
                           int w_is_device, double* grad, double* loss, smg_stats* stats);

/* Matrix-free projections with the PE matrix of rotation `rotation`: each
 * element is computed as smg_pe_forward computes it (FP32 value) and used at
 * once, never stored.  Forward projection y[d] = sum_v A[d][v] x[v] (x: V
 * floats, y: N floats); back-projection x[v] = sum_d A[d][v] y[d] (bars with
 * y[d] = 0 are skipped).  FP64 sums in a fixed order (bitwise reproducible).
 * No counterpart in the reference, which writes the matrix to a file for a
 * reconstruction program to read (PEGen_RayTracing_CircularHole/
 * PEGen_CircularHole.cpp:141-147). */
smg_status smg_pe_project(smg_ctx* ctx, int rotation, smg_precision prec, const float* x,
                          int x_is_device, float* y, int y_is_device, smg_stats* stats);
smg_status smg_pe_backproject(smg_ctx* ctx, int rotation, smg_precision prec, const float* y,
                              int y_is_device, float* x, int x_is_device, smg_stats* stats);

/* ---- geometry bitmaps (the scatter kernels' attenuation culling) ------------
 * Crystal bitmap: N(N+1)/2 * N bits; collimator bitmap: H0 * N * N bits,
 * packed little-endian in 32-bit words, the reference's cache-file format
 * (ScatterGen_RayTracing_CircularHole/scatter.cu:2053-2063, :2233-2243). */
int64_t smg_bitmap_crystal_words(int n_det);
int64_t smg_bitmap_collimator_words(int n_holes_l0, int n_det);
/* Compute the bitmap on the GPU, keep it in the context, and copy it to
 * words (may be NULL) of n_words words. */
smg_status smg_bitmap_crystal(smg_ctx* ctx, uint32_t* words, int64_t n_words, int words_is_device);
smg_status smg_bitmap_collimator(smg_ctx* ctx, uint32_t* words, int64_t n_words,
                                 int words_is_device);
/* Install a bitmap read from a cache file instead of computing it. */
smg_status smg_load_bitmap_crystal(smg_ctx* ctx, const uint32_t* words, int64_t n_words);
smg_status smg_load_bitmap_collimator(smg_ctx* ctx, const uint32_t* words, int64_t n_words);

/* ---- primary Compton scatter (replaces the kernels behind scatter()) --------
 * Crystal scatter: pe is the N x V PE block used as scatter weights
 * (ScatterGen_RayTracing_CircularHole/scatter.cu:1077-1437).  Uses the
 * crystal bitmap of the context (computed on first use if none is set). */
smg_status smg_scatter_crystal(smg_ctx* ctx, int rotation, smg_precision prec, const float* pe,
                               int pe_is_device, float* dst, int dst_is_device, smg_stats* stats);
/* Collimator scatter at the layer-0 hole centres (scatter.cu:1544-1936).
 * Uses the collimator bitmap of the context (computed on first use). */
smg_status smg_scatter_collimator(smg_ctx* ctx, int rotation, smg_coll_variant variant,
                                  smg_precision prec, float* dst, int dst_is_device,
                                  smg_stats* stats);
/* The scatter block of the parts in `parts` (smg_scatter_part bits: crystal
 * scatter driven by pe_drive, and one collimator variant), summed on the GPU,
 * into scatter_dst, and pe_add + scatter into sum_dst: the blocks the reference
 * drivers form on the host (scatter() of the combined generator adds crystal
 * and collimator scatter, ScatterGen_RayTracing_CircularHole/scatter.cu:
 * 2354-2357; every driver writes PE + scatter, ScatterGen_CircularHole.cpp:
 * 211-230, ScatterGen_Collimator.cpp and ScatterGen_Crystal.cpp likewise).
 * Either output may be NULL (not both); pe_drive and pe_add may be the same
 * block; both inputs are host or both device memory (pe_is_device), both
 * outputs likewise (dst_is_device).  Same values as the separate calls: the
 * parts are added in float32 in the order crystal, collimator, then PE. */
smg_status smg_scatter_sum(smg_ctx* ctx, int rotation, unsigned parts, smg_precision prec,
                           const float* pe_drive, const float* pe_add, int pe_is_device,
                           float* scatter_dst, float* sum_dst, int dst_is_device, smg_stats* stats);

/* Collimator scatter with scatter points on a grid over the tungsten of the
 * layers in layer_mask (bit m: layer m; 1 = the layer-0 plate), pitch_mm,
 * cells of volume dv; points inside a hole of their layer left out.  No
 * counterpart in the reference, whose collimator scatterers are the layer-0
 * hole centres (ScatterGen_RayTracing_CircularHole/scatter.cu:1702-1716).
 * Each point weighs dv mu_c / (4 pi L^2) exp(-collimator material on the way
 * from the voxel), mu_c of its layer; the second leg passes every other bar
 * and the collimator material at E' (mu(E) tables or power laws); the rest
 * is the reference's scatter term.  n_points (optional) receives the number
 * of grid points. */
smg_status smg_scatter_collimator_grid(smg_ctx* ctx, int rotation, double pitch_mm, unsigned layer_mask,
                                       smg_precision prec, float* dst, int dst_is_device,
                                       int64_t* n_points, smg_stats* stats);

/* Crystal scatter with scatter points at the centres of a gx x gy x gz grid
 * of cells in every bar, for the detector rows `rows` (n_rows of them; dst
 * row r is bar rows[r]) or for all bars when rows is NULL.  No counterpart in
 * the reference, whose crystal scatterer is each bar's centre weighted by the
 * PE matrix (ScatterGen_Crystal/scatter.cu:1061-1421).  Each point weighs
 * dv mu_c / (4 pi L^2) exp(-att), att being the photoelectric ray's
 * attenuation to the point; the second leg passes the bars flagged in the
 * crystal bitmap for the bar pair and leaves the scatter bar; the rest is the
 * reference's scatter term. */
smg_status smg_scatter_crystal_grid(smg_ctx* ctx, int rotation, int gx, int gy, int gz,
                                    const int* rows, int n_rows, smg_precision prec, float* dst,
                                    int dst_is_device, smg_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* SMG_H */
