/*
 * FP64 host oracle for the SPECT/PET system-matrix generator.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load this library, and only as the checker.
 * The shipped library never links or calls it.
 *
 * Every function restates, in double precision, the formula of the
 * reference kernel it cites (paths relative to the reference repository).
 * The geometry parameters come in as the same float32 arrays the GPU path
 * reads, converted to double without rounding.
 *
 * Gradient layout (adjoint): [mu of each collimator layer (Lc)]
 *                            [mu_total of each bar (N)] [mu_pe of each bar (N)].
 */
#ifndef SMG_ORACLE_H
#define SMG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smg_o_geom smg_o_geom;

/* Parse the four parameter arrays (80000/80000/100/100 floats).  The arrays
 * are referenced, not copied, and must outlive the handle.  Returns NULL if
 * a count is out of range. */
smg_o_geom* smg_o_geom_new(const float* coll, const float* det, const float* img,
                           const float* phys);
void smg_o_geom_free(smg_o_geom* g);
int smg_o_n_det(const smg_o_geom* g);
int smg_o_n_vox(const smg_o_geom* g);
int smg_o_n_holes_l0(const smg_o_geom* g);
int smg_o_n_layers(const smg_o_geom* g);

/* ---- geometry primitives ------------------------------------------------ */
double smg_o_box_ray(const double a[3], const double b[3], const double lo[3],
                     const double hi[3]);
double smg_o_box_ray_inside(const double a[3], const double b[3], const double lo[3],
                            const double hi[3]);
double smg_o_cylinder_ray(const double a[3], const double b[3], double xc, double y1,
                          double y2, double zc, double radius);
/* Elliptic-cylinder chord (PE:801-947, defined but never called by the
 * reference): a and b turned by theta about the axis, then the ellipse
 * ((x-xc)/x_rho)^2 + ((z-zc)/z_rho)^2 <= 1 between y1 and y2. */
double smg_o_ellcyl_ray(const double a[3], const double b[3], double xc, double y1, double y2,
                        double zc, double x_rho, double z_rho, double theta);
/* Elliptical holes for the PE element: x_rho, z_rho, theta per hole (the
 * array is not copied; NULL = all circles).  Culling uses the bounding radius
 * max(x_rho, z_rho). */
void smg_o_set_hole_shapes(smg_o_geom* g, const float* shapes);

/* ---- Compton physics helpers ------------------------------------------- */
double smg_o_kn(double theta, double e0);
double smg_o_kn_sum(double e0, double lo, double hi, double step);
double smg_o_cone_angle(const double src[3], const double a[3], const double det[3]);
double smg_o_scatter_energy(double theta, double e0);
double smg_o_gauss_window(double e, double res, double lo, double hi);

/* ---- photoelectric matrix ---------------------------------------------- */
double smg_o_pe_element(const smg_o_geom* g, int d, int v, int rot, int sx, int sy, int sz);
/* rows[i] = detector index; out is [nrows][V]; OpenMP over rows. */
void smg_o_pe_rows(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* rows,
                   int nrows, double* out, int nthreads);
/* Elements listed as (d, v) pairs. */
void smg_o_pe_pairs(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* d,
                    const int* v, int n, double* out, int nthreads);
/* Culling counts of one element (crystals kept, holes kept over all layers). */
void smg_o_pe_cull_counts(const smg_o_geom* g, int d, int v, int rot, int* n_cryst,
                          int* n_holes);

/* Algorithmic arithmetic operations per sub-cell ray of the reference PE
 * formula (branches taken; culling scans and loop-invariant trig excluded),
 * averaged over the listed elements. */
/* Algorithmic operations per scatter term (row, voxel, scatter point,
 * detection cell) of the reference's scatter element body on the branches
 * taken; kind 0 crystal, 1 collimator (hole centres), 2 collimator grid,
 * 3 crystal grid.  Counted on the calling thread; *terms = terms counted. */
double smg_o_scatter_ops_per_term(const smg_o_geom* g, const uint32_t* bitmap, int kind,
                                  double pitch, unsigned layer_mask, int gx, int gy, int gz, int rot,
                                  int sx, int sy, int sz, const int* d, const int* v, int n,
                                  double* terms);
double smg_o_pe_ops_per_ray(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* d,
                            const int* v, int n);

/* dL/dmu for L = sum_dv w[d][v] A[d][v]; grad has Lc + 2N entries. */
void smg_o_pe_grad(const smg_o_geom* g, int rot, int sx, int sy, int sz, const double* w,
                   double* grad, double* loss, int nthreads);

/* ---- geometry bitmaps --------------------------------------------------- */
int64_t smg_o_bitmap_crystal_words(int n_det);
int64_t smg_o_bitmap_collimator_words(int n_holes_l0, int n_det);
void smg_o_bitmap_crystal(const smg_o_geom* g, uint32_t* words, int nthreads);
void smg_o_bitmap_collimator(const smg_o_geom* g, uint32_t* words, int nthreads);
/* The predicates behind single bits (i < j for the crystal bitmap). */
int smg_o_crystal_pred(const smg_o_geom* g, int i, int j, int k);
int smg_o_coll_pred(const smg_o_geom* g, int i, int j, int k);
int smg_o_bit_crystal(const uint32_t* words, int n_det, int i, int j, int k);
int smg_o_bit_collimator(const uint32_t* words, int n_det, int i, int j, int k);

/* ---- mu(E) tables (extension, no reference counterpart) ------------------ */
/* Node energies (keV, increasing) and per-process values: crystal pe, crystal
 * Compton, collimator pe, collimator Compton.  Linear in (log E, log mu).
 * n = 0 restores the reference's power laws.  Returns 0, or -1 on bad input. */
int smg_o_set_mu_tables(smg_o_geom* g, int n, const double* energy, const double* crystal_pe,
                        const double* crystal_c, const double* coll_pe, const double* coll_c);
/* mu(e)/mu(e0) of process `which` (0..3 as above). */
double smg_o_mu_ratio(const smg_o_geom* g, int which, double e, double e0);

/* ---- primary Compton scatter -------------------------------------------- */
/* pe is the [N][V] PE slice of the voxel grid; sx,sy,sz = detection-cell
 * subsampling (the reference uses 1,1,1). */
double smg_o_crystal_scatter_element(const smg_o_geom* g, const float* pe,
                                     const uint32_t* bitmap, int d, int v, int rot, int sx,
                                     int sy, int sz);
/* variant 0: combined generator (r = 0.1 mm); 1: standalone (r = 0.2 mm). */
double smg_o_collimator_scatter_element(const smg_o_geom* g, const uint32_t* bitmap, int d,
                                        int v, int rot, int variant, int sx, int sy, int sz);
void smg_o_crystal_scatter_pairs(const smg_o_geom* g, const float* pe, const uint32_t* bitmap,
                                 int rot, int sx, int sy, int sz, const int* d, const int* v,
                                 int n, double* out, int nthreads);
void smg_o_collimator_scatter_pairs(const smg_o_geom* g, const uint32_t* bitmap, int rot,
                                    int variant, int sx, int sy, int sz, const int* d,
                                    const int* v, int n, double* out, int nthreads);

/* ---- collimator scatter from grid points (extension, no counterpart) ---- */
double smg_o_seg_box(const double a[3], const double b[3], const double lo[3], const double hi[3]);
double smg_o_seg_cyl(const double a[3], const double b[3], double xc, double y1, double y2,
                     double zc, double r);
double smg_o_seg_ell(const double a[3], const double b[3], double xc, double y1, double y2,
                     double zc, double x_rho, double z_rho, double theta);
int64_t smg_o_coll_grid_points(const smg_o_geom* g, double pitch, unsigned layer_mask, double* xyzw,
                               int64_t max);
double smg_o_coll_grid_weight(const smg_o_geom* g, const double p[3], const double s[3], double w);
void smg_o_coll_grid_scatter_pairs(const smg_o_geom* g, double pitch, unsigned layer_mask, int rot,
                                   int sx, int sy, int sz, const int* d, const int* v, int n,
                                   double* out, int nthreads);

/* ---- crystal scatter from grid points (extension, no counterpart) -------- */
double smg_o_crystal_grid_weight(const smg_o_geom* g, int b, int v, int rot, int gx, int gy,
                                 int gz, int cell);
void smg_o_crystal_grid_scatter_pairs(const smg_o_geom* g, const uint32_t* bitmap, int rot,
                                      int gx, int gy, int gz, int sx, int sy, int sz,
                                      const int* d, const int* v, int n, double* out,
                                      int nthreads);

int smg_o_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
