/*
 * FP64 host oracle: a restatement of the reference kernel formulas in double
 * precision.  TEST INFRASTRUCTURE ONLY (tests, smoke check, CPU baseline);
 * the shipped library never calls it.
 *
 * Reference files cited (relative to the reference repository):
 *   PE   = PEGen_RayTracing_CircularHole/PESysMatGen.cu
 *   SC   = ScatterGen_RayTracing_CircularHole/scatter.cu   (combined program)
 *   SCC  = ScatterGen_Collimator/scatter.cu                (standalone variant)
 */
#include "smg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants keep the value they have in the reference: float literals are
 * widened from float, double literals stay double. */
#define SLAB_EPS ((double)0.001f) /* PE:26, PE:326 (float eps) */
#define INSIDE_DY 0.001           /* PE:321 (double literal) */
#define CYL_DY 0.001              /* PE:695 */
#define CYL_TEPS 0.0001           /* PE:732-733 */
#define CYL_DEPS 0.00001          /* PE:745 */
#define KN_STEP ((double)0.01f)   /* SC:1338, 0.01 passed as float */
#define KN_PI ((double)(float)M_PI) /* SC:1169, M_PI passed as float */
#define FWHM_SIGMA ((double)2.35482f) /* SC:1026, SC:1263 */

/* ---- algorithmic op counter ---------------------------------------------
 * When enabled (smg_o_pe_ops_per_ray), the element code adds the arithmetic
 * operations (+ - * / sqrt exp; comparisons excluded) of the reference
 * formula on the branch it takes.  Loop-invariant trig of the bar rotation
 * and the culling scans are not counted. */
static _Thread_local int g_count_ops = 0;
static _Thread_local double g_ops = 0.0;
static _Thread_local double g_terms = 0.0; /* scatter terms evaluated while counting */
#define OPS(n)                          \
    do {                                \
        if (g_count_ops) g_ops += (n);  \
    } while (0)

#define SMG_O_MAX_NODES 256

struct smg_o_geom {
    const float* coll;
    const float* det;
    const float* img;
    const float* phys;
    int n_layers, n_det, nx, ny, nz, n_vox, n_holes_total, n_holes_l0;
    int holes_in_layer[16];
    double F;
    /* mu(E) tables: log E of the nodes, log mu per process (0 crystal pe,
     * 1 crystal compton, 2 collimator pe, 3 collimator compton) */
    int n_tab;
    double tab_le[SMG_O_MAX_NODES];
    double tab_lmu[4][SMG_O_MAX_NODES];
    /* elliptical holes: x_rho, z_rho, theta per hole (NULL: all circles) */
    const float* hole_shapes;
};

smg_o_geom* smg_o_geom_new(const float* coll, const float* det, const float* img,
                           const float* phys) {
    smg_o_geom* g = (smg_o_geom*)calloc(1, sizeof(smg_o_geom));
    if (!g) return NULL;
    g->coll = coll;
    g->det = det;
    g->img = img;
    g->phys = phys;
    /* Integer fields are stored as floats and read with floor(x + small eps)
     * (PE:961-964, PE:1377-1383). */
    g->n_layers = (int)floor((double)coll[0] + 1e-6);
    g->n_det = (int)floor((double)det[0] + 1e-4);
    g->nx = (int)floor((double)img[0] + 1e-3);
    g->ny = (int)floor((double)img[1] + 1e-3);
    g->nz = (int)floor((double)img[2] + 1e-3);
    g->n_vox = g->nx * g->ny * g->nz;
    g->F = img[11];
    if (g->n_layers < 0 || g->n_layers > 9 || g->n_det < 0 || g->n_det > 6666 ||
        g->nx <= 0 || g->ny <= 0 || g->nz <= 0) {
        free(g);
        return NULL;
    }
    g->n_holes_total = 0;
    for (int m = 0; m < g->n_layers; ++m) {
        g->holes_in_layer[m] = (int)floor((double)coll[(m + 1) * 10]); /* PE:980-982 */
        g->n_holes_total += g->holes_in_layer[m];
    }
    g->n_holes_l0 = (int)floor((double)coll[10] + 1e-6); /* SC:1555 */
    if (g->n_holes_total > (80000 - 100) / 9) {
        free(g);
        return NULL;
    }
    return g;
}

void smg_o_geom_free(smg_o_geom* g) { free(g); }
int smg_o_n_det(const smg_o_geom* g) { return g->n_det; }
int smg_o_n_vox(const smg_o_geom* g) { return g->n_vox; }
int smg_o_n_holes_l0(const smg_o_geom* g) { return g->n_holes_l0; }
int smg_o_n_layers(const smg_o_geom* g) { return g->n_layers; }

int smg_o_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================== */
/* Geometry primitives                                                      */
/* ======================================================================== */

/* Interval of the ray parameter t (distance from a along the unit direction
 * of b - a, scaled by len) inside the slabs of the listed axes.  The
 * reference intersects the per-axis intervals one after the other with an
 * early exit on an empty intersection (PE:77-111); the running intersection
 * is empty exactly when the final one is, and an empty final interval fails
 * the (tmax - tmin) < eps test, so plain max/min of the per-axis bounds gives
 * the same result. */
static void slab_interval(const double a[3], const double b[3], const double lo[3],
                          const double hi[3], const int* axes, int naxes, double len,
                          double* tmin_out, double* tmax_out) {
    double tmin = -INFINITY, tmax = INFINITY;
    for (int i = 0; i < naxes; ++i) {
        int k = axes[i];
        double inv = len / (b[k] - a[k]);
        double t_lo, t_hi;
        if (inv < 0) {
            t_lo = (hi[k] - a[k]) * inv;
            t_hi = (lo[k] - a[k]) * inv;
        } else {
            t_lo = (lo[k] - a[k]) * inv;
            t_hi = (hi[k] - a[k]) * inv;
        }
        if (t_lo > tmin) tmin = t_lo;
        if (t_hi < tmax) tmax = t_hi;
    }
    *tmin_out = tmin;
    *tmax_out = tmax;
}

/* Chord of the segment a -> b through the axis-aligned box [lo, hi]
 * (PE:20-312).  A direction component below eps = 0.001 is treated as zero:
 * with two such components the segment must run past both faces of the box
 * along the third axis (PE:35-66); with one, the segment must lie strictly
 * between that axis's faces and the slab test runs on the other two axes
 * with the 2-D length (PE:69-232); otherwise the 3-D slab test (PE:236-312).
 * A chord counts only if both slab parameters lie in (eps, |b-a|) and the
 * chord is at least eps long (PE:297-304). */
double smg_o_box_ray(const double a[3], const double b[3], const double lo[3],
                     const double hi[3]) {
    const double eps = SLAB_EPS;
    int small[3];
    for (int k = 0; k < 3; ++k) small[k] = fabs(b[k] - a[k]) < eps;
    int nsmall = small[0] + small[1] + small[2];
    OPS(3); /* component differences */
    if (nsmall == 3) return 0.0;
    if (nsmall == 2) {
        OPS(1);
        int k = !small[0] ? 0 : (!small[1] ? 1 : 2); /* the moving axis */
        int i = (k + 1) % 3, j = (k + 2) % 3;
        if (a[i] >= lo[i] && a[i] <= hi[i] && a[j] >= lo[j] && a[j] <= hi[j]) {
            if ((a[k] <= lo[k] && b[k] >= hi[k]) || (b[k] <= lo[k] && a[k] >= hi[k]))
                return fabs(hi[k] - lo[k]);
        }
        return 0.0;
    }
    int axes[3], naxes = 0;
    double len2 = 0.0;
    if (nsmall == 1) {
        int s = small[0] ? 0 : (small[1] ? 1 : 2);
        if (a[s] >= hi[s] || a[s] <= lo[s]) return 0.0;
        for (int k = 0; k < 3; ++k)
            if (k != s) {
                axes[naxes++] = k;
                len2 += (b[k] - a[k]) * (b[k] - a[k]);
            }
    } else {
        for (int k = 0; k < 3; ++k) {
            axes[naxes++] = k;
            len2 += (b[k] - a[k]) * (b[k] - a[k]);
        }
    }
    double len = sqrt(len2);
    double tmin, tmax;
    slab_interval(a, b, lo, hi, axes, naxes, len, &tmin, &tmax);
    /* length (naxes mul, naxes-1 add, sqrt), naxes div, 2 naxes (sub, mul), chord */
    OPS(naxes + (naxes - 1) + 1 + naxes + 4 * naxes + 1);
    if (tmax - tmin < eps) return 0.0;
    if (tmin >= len || tmax >= len) return 0.0;
    if (tmin <= eps || tmax <= eps) return 0.0;
    return tmax - tmin;
}

/* Chord of the segment a -> b through the box when b may lie inside it
 * (PE:314-686): the chord is clipped at b.  A segment whose y component is
 * below 0.001 gives 0 (PE:321-324); the branches for components below eps
 * follow the same order as in the chord above.  Clipping rules: PE:415-434. */
double smg_o_box_ray_inside(const double a[3], const double b[3], const double lo[3],
                            const double hi[3]) {
    const double eps = SLAB_EPS;
    OPS(3);
    if (fabs(b[1] - a[1]) < INSIDE_DY) return 0.0;
    int small[3];
    for (int k = 0; k < 3; ++k) small[k] = fabs(b[k] - a[k]) < eps;
    if (small[0] && small[1] && small[2]) return 0.0;
    /* two components below eps: the segment runs along the third axis and
     * the chord is measured from the entry face to b (PE:333-382) */
    static const int pairs[3][3] = {{0, 1, 2}, {2, 1, 0}, {0, 2, 1}}; /* (i, j, moving k) */
    for (int pi = 0; pi < 3; ++pi) {
        int i = pairs[pi][0], j = pairs[pi][1], k = pairs[pi][2];
        if (small[i] && small[j]) {
            if (a[i] >= lo[i] && a[i] <= hi[i] && a[j] >= lo[j] && a[j] <= hi[j]) {
                if (a[k] < b[k]) return b[k] - lo[k];
                if (a[k] > b[k]) return hi[k] - b[k];
            } else {
                return 0.0;
            }
        }
    }
    int axes[3], naxes = 0;
    double len2 = 0.0;
    int s = small[0] ? 0 : (small[1] ? 1 : (small[2] ? 2 : -1));
    if (s >= 0) {
        if (a[s] >= hi[s] || a[s] <= lo[s]) return 0.0;
        for (int k = 0; k < 3; ++k)
            if (k != s) {
                axes[naxes++] = k;
                len2 += (b[k] - a[k]) * (b[k] - a[k]);
            }
    } else {
        for (int k = 0; k < 3; ++k) {
            axes[naxes++] = k;
            len2 += (b[k] - a[k]) * (b[k] - a[k]);
        }
    }
    double len = sqrt(len2);
    double tmin, tmax;
    slab_interval(a, b, lo, hi, axes, naxes, len, &tmin, &tmax);
    OPS(naxes + (naxes - 1) + 1 + naxes + 4 * naxes + 1);
    if (tmax - tmin < eps) return 0.0;
    if (tmin >= len) return 0.0;
    if (tmax <= eps) return 0.0;
    if (tmax >= len && tmin > eps) return len - tmin;
    if (tmin <= eps && tmax <= len) return tmax;
    if (tmin <= eps && tmax >= len) return len;
    return tmax - tmin;
}

/* Chord of the segment a -> b through the finite cylinder of radius R whose
 * axis is parallel to y at (xc, zc) and which spans y1..y2 (PE:688-799).
 * The plane points where the line meets y = y1 and y = y2 decide the case:
 * both inside the circle -> the slab chord; otherwise the quadratic of the
 * infinite cylinder, cut by whichever plane point is inside. */
double smg_o_cylinder_ray(const double a[3], const double b[3], double xc, double y1,
                          double y2, double zc, double radius) {
    OPS(1);
    if (fabs(y1 - y2) < CYL_DY) return 0.0;
    /* length and unit direction 12, plane points 16, slab 4, circle tests 12 */
    OPS(44);
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    double len = sqrt(dx * dx + dy * dy + dz * dz);
    double kx = dx / len, ky = dy / len, kz = dz / len;
    double xl = a[0] + kx / ky * (y1 - a[1]);
    double xr = a[0] + kx / ky * (y2 - a[1]);
    double zl = a[2] + kz / ky * (y1 - a[1]);
    double zr = a[2] + kz / ky * (y2 - a[1]);
    double tmin = (y1 - a[1]) / ky;
    double tmax = (y2 - a[1]) / ky;
    double r2 = radius * radius;
    int in_l = (xl - xc) * (xl - xc) + (zl - zc) * (zl - zc) <= r2;
    int in_r = (xr - xc) * (xr - xc) + (zr - zc) * (zr - zc) <= r2;
    if (in_l && in_r) {
        OPS(1);
        if (tmin <= CYL_TEPS || tmax <= CYL_TEPS) return 0.0;
        if (tmin >= len || tmax >= len) return 0.0;
        return fabs(tmax - tmin);
    }
    OPS(12); /* offsets and discriminant */
    double x0 = a[0] - xc, z0 = a[2] - zc;
    double qa = kx * kx + kz * kz;
    double cr = kx * z0 - kz * x0;
    double disc = qa * r2 - cr * cr;
    if (disc <= CYL_DEPS) return 0.0;
    OPS(20); /* sqrt, two roots, chord */
    double sq = sqrt(disc);
    double qb = kx * x0 + kz * z0;
    double t1 = (-qb - sq) / qa;
    double t2 = (-qb + sq) / qa;
    if (!in_l && !in_r) {
        if (t1 >= tmin && t2 <= tmax) {
            if (t2 <= CYL_TEPS || t1 <= CYL_TEPS) return 0.0;
            if (t2 >= len || t1 >= len) return 0.0;
            return t2 - t1;
        }
        return 0.0;
    }
    if (in_l) { /* enters through the y1 plane, leaves through the wall */
        if (t2 <= CYL_TEPS || tmin <= CYL_TEPS) return 0.0;
        if (t2 >= len || tmin >= len) return 0.0;
        return t2 - tmin;
    }
    /* enters through the wall, leaves through the y2 plane */
    if (tmax <= CYL_TEPS || t1 <= CYL_TEPS) return 0.0;
    if (tmax >= len || t1 >= len) return 0.0;
    return tmax - t1;
}

/* Chord of the segment a -> b through the finite elliptic cylinder
 * ((x'-xc)/x_rho)^2 + ((z'-zc)/z_rho)^2 <= 1, y1 <= y <= y2, with a and b
 * first turned by theta about the axis (xc, zc): x' = xc + dx cos + dz sin,
 * z' = zc - dx sin + dz cos (PE:801-947; the reference defines this routine
 * but never calls it).  After the turn it follows the circular chord's cases
 * with the ellipse's quadratic a = kx^2 z_rho^2 + kz^2 x_rho^2,
 * b = kx x0 z_rho^2 + kz z0 x_rho^2, c = z_rho^2 x0^2 + x_rho^2 z0^2 -
 * x_rho^2 z_rho^2 and the cut b^2 - a c <= 1e-5 (PE:887-893). */
double smg_o_ellcyl_ray(const double a[3], const double b[3], double xc, double y1, double y2,
                        double zc, double x_rho, double z_rho, double theta) {
    double cs = cos(theta), sn = sin(theta);
    double axr = a[0] - xc, azr = a[2] - zc, bxr = b[0] - xc, bzr = b[2] - zc;
    double x_in = xc + axr * cs + azr * sn, z_in = zc - axr * sn + azr * cs;
    double x_out = xc + bxr * cs + bzr * sn, z_out = zc - bxr * sn + bzr * cs;
    double y_in = a[1], y_out = b[1];
    if (fabs(y1 - y2) < CYL_DY) return 0.0;
    double len = sqrt((x_out - x_in) * (x_out - x_in) + (y_out - y_in) * (y_out - y_in) +
                      (z_out - z_in) * (z_out - z_in));
    double kx = (x_out - x_in) / len, ky = (y_out - y_in) / len, kz = (z_out - z_in) / len;
    double xl = x_in + kx / ky * (y1 - y_in), xr = x_in + kx / ky * (y2 - y_in);
    double zl = z_in + kz / ky * (y1 - y_in), zr = z_in + kz / ky * (y2 - y_in);
    double tmin = (y1 - y_in) / ky, tmax = (y2 - y_in) / ky;
    int in_l = (xl - xc) * (xl - xc) / x_rho / x_rho + (zl - zc) * (zl - zc) / z_rho / z_rho <= 1;
    int in_r = (xr - xc) * (xr - xc) / x_rho / x_rho + (zr - zc) * (zr - zc) / z_rho / z_rho <= 1;
    if (in_l && in_r) {
        if (tmin <= CYL_TEPS || tmax <= CYL_TEPS) return 0.0;
        if (tmin >= len || tmax >= len) return 0.0;
        return fabs(tmax - tmin);
    }
    double x0 = x_in - xc, z0 = z_in - zc;
    double qa = kx * kx * z_rho * z_rho + kz * kz * x_rho * x_rho;
    double qb = kx * x0 * z_rho * z_rho + kz * z0 * x_rho * x_rho;
    double qc = z_rho * z_rho * x0 * x0 + x_rho * x_rho * z0 * z0 - x_rho * x_rho * z_rho * z_rho;
    double disc = qb * qb - qa * qc;
    if (disc <= CYL_DEPS) return 0.0;
    double t1 = (-qb - sqrt(disc)) / qa;
    double t2 = (-qb + sqrt(disc)) / qa;
    if (!in_l && !in_r) {
        if (t1 >= tmin && t2 <= tmax) {
            if (t2 <= CYL_TEPS || t1 <= CYL_TEPS) return 0.0;
            if (t2 >= len || t1 >= len) return 0.0;
            return t2 - t1;
        }
        return 0.0;
    }
    if (in_l) {
        if (t2 <= CYL_TEPS || tmin <= CYL_TEPS) return 0.0;
        if (t2 >= len || tmin >= len) return 0.0;
        return t2 - tmin;
    }
    if (tmax <= CYL_TEPS || t1 <= CYL_TEPS) return 0.0;
    if (tmax >= len || t1 >= len) return 0.0;
    return tmax - t1;
}

void smg_o_set_hole_shapes(smg_o_geom* g, const float* shapes) { g->hole_shapes = shapes; }

/* Elliptical shape of hole h, or 0 for a circle. */
static int hole_ellipse(const smg_o_geom* g, int h, double* xr, double* zr, double* th) {
    if (!g->hole_shapes || !(g->hole_shapes[3 * h] > 0.0f)) return 0;
    *xr = g->hole_shapes[3 * h];
    *zr = g->hole_shapes[3 * h + 1];
    *th = g->hole_shapes[3 * h + 2];
    return 1;
}

/* ======================================================================== */
/* Compton physics helpers (SC:961-1038)                                    */
/* ======================================================================== */

/* Klein-Nishina dsigma/dOmega without the r_e^2/2 prefactor, in the
 * factorised form of SC:962-977: with c = cos(theta), a = E0/511,
 * f1 = a(1-c), f2 = 1+c^2:  f2/(1+f1)^2 * (1 + f1^2/(f2 (1+f1))). */
double smg_o_kn(double theta, double e0) {
    OPS(16); /* cos, alpha, f1, f2, the quotient as written (1 + f1 three times) */
    double c = cos(theta);
    double alpha = e0 / 511.0;
    double f1 = alpha * (1.0 - c);
    double f2 = 1.0 + c * c;
    return f2 / (1.0 + f1) / (1.0 + f1) * (1.0 + f1 * f1 / f2 / (1.0 + f1));
}

/* Inclusive rectangle rule of KN(theta) sin(theta) on lo, lo+step, ...,
 * lo + n*step with n = (int)((hi - lo)/step) truncated toward zero
 * (SC:980-993). */
double smg_o_kn_sum(double e0, double lo, double hi, double step) {
    int n = (int)((hi - lo) / step);
    OPS(2);
    double s = 0.0;
    for (int i = 0; i <= n; ++i) {
        OPS(6); /* node angle, sin, two products, sum */
        double th = lo + i * step;
        s += smg_o_kn(th, e0) * sin(th) * step;
    }
    return s;
}

/* Angle between src -> a and a -> det, acos of the clamped cosine
 * (SC:996-1013). */
double smg_o_cone_angle(const double src[3], const double a[3], const double det[3]) {
    OPS(26); /* 6 differences, dot, two norms with sqrt, quotient, acos */
    double u[3] = {a[0] - src[0], a[1] - src[1], a[2] - src[2]};
    double w[3] = {det[0] - a[0], det[1] - a[1], det[2] - a[2]};
    double dot = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    double nu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double c = dot / (nu * nw);
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    return acos(c);
}

/* E' = E0 / (1 + (E0/511)(1 - cos theta))  (SC:1016-1023). */
double smg_o_scatter_energy(double theta, double e0) {
    OPS(6);
    return e0 / (1.0 + (e0 / 511.0) * (1.0 - cos(theta)));
}

/* Probability that a Gaussian of FWHM res*e centred on e falls in [lo, hi]
 * (SC:1025-1033). */
double smg_o_gauss_window(double e, double res, double lo, double hi) {
    OPS(12); /* sigma, sqrt(2) sigma, two arguments, two erf, difference, half */
    double sigma = res / FWHM_SIGMA * e;
    double s2 = sigma * sqrt(2.0);
    double z1 = (lo - e) / s2;
    double z2 = (hi - e) / s2;
    return 0.5 * (erf(z2) - erf(z1));
}

/* ======================================================================== */
/* Shared geometry of the kernels                                           */
/* ======================================================================== */

typedef struct {
    double c[3];       /* centre, y includes F */
    double w, t, h;    /* width (x), thickness (y), height (z) */
    double mut, mupe, muc, res, beta;
} bar_t;

static bar_t get_bar(const smg_o_geom* g, int k) {
    const float* d = g->det + 12 * k;
    bar_t b;
    b.c[0] = d[1];
    b.c[1] = (double)d[2] + g->F;
    b.c[2] = d[3];
    b.w = d[4];
    b.t = d[5];
    b.h = d[6];
    b.mut = d[7];
    b.mupe = d[8];
    b.muc = d[9];
    b.res = d[10];
    b.beta = d[11];
    return b;
}

/* World point -> bar frame: rotate by -beta about y around the bar centre
 * (PE:1071-1073, PE:1309-1311). */
static void to_bar_frame(const bar_t* b, const double u[3], double out[3]) {
    OPS(9); /* rotation trig of the bar is loop-invariant: not counted */
    double cb = cos(-b->beta), sb = sin(-b->beta);
    double x = u[0] - b->c[0], z = u[2] - b->c[2];
    out[0] = x * cb - z * sb;
    out[1] = u[1] - b->c[1];
    out[2] = x * sb + z * cb;
}

/* Bar frame -> world (PE:1211-1217). */
static void from_bar_frame(const bar_t* b, const double s[3], double out[3]) {
    OPS(9);
    double cb = cos(b->beta), sb = sin(b->beta);
    out[0] = b->c[0] + s[0] * cb - s[2] * sb;
    out[1] = b->c[1] + s[1];
    out[2] = b->c[2] + s[0] * sb + s[2] * cb;
}

/* Voxel centre, shifted, rotated about z by rot * dphi (PE:1004-1053). */
static void voxel_pos(const smg_o_geom* g, int v, int rot, double p[3]) {
    int nxy = g->nx * g->ny;
    int iz = v / nxy, rem = v % nxy, iy = rem / g->nx, ix = rem % g->nx;
    double x = (ix - g->nx / 2.0 + 0.5) * (double)g->img[3] + (double)g->img[8];
    double y = (iy - g->ny / 2.0 + 0.5) * (double)g->img[4] + (double)g->img[9];
    double z = (iz - g->nz / 2.0 + 0.5) * (double)g->img[5] + (double)g->img[10];
    double phi = (double)rot * (double)g->img[7];
    p[0] = x * cos(phi) - y * sin(phi);
    p[1] = x * sin(phi) + y * cos(phi);
    p[2] = z;
}

/* Sub-cell (a, b, c) of an (sx, sy, sz) split of bar b: centre and box in the
 * bar frame (PE:1207-1209, PE:1329-1336). */
static void cell_of(const bar_t* b, int a, int bb, int c, int sx, int sy, int sz, double ctr[3],
                    double lo[3], double hi[3]) {
    ctr[0] = -b->w / 2.0 + (a + 0.5) / sx * b->w;
    ctr[1] = -b->t / 2.0 + (bb + 0.5) / sy * b->t;
    ctr[2] = -b->h / 2.0 + (c + 0.5) / sz * b->h;
    lo[0] = ((double)a / sx - 0.5) * b->w;
    hi[0] = (((double)a + 1.0) / sx - 0.5) * b->w;
    lo[1] = ((double)bb / sy - 0.5) * b->t;
    hi[1] = (((double)bb + 1.0) / sy - 0.5) * b->t;
    lo[2] = ((double)c / sz - 0.5) * b->h;
    hi[2] = (((double)c + 1.0) / sz - 0.5) * b->h;
}

static double dist3(const double a[3], const double b[3]) {
    OPS(9);
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* ======================================================================== */
/* Photoelectric element (PE:951-1364)                                      */
/* ======================================================================== */

typedef struct {
    int n_cryst;
    int* cryst;          /* culled bars */
    int n_holes[16];
    int* holes[16];      /* culled holes per layer (global hole index) */
} cull_t;

/* Culling by bounding spheres projected on the unit sphere around the voxel
 * (PE:1097-1186): bar k (k != d) is kept when |u_k - u_d| <= r_k/D_k + r_d/D_d
 * with r = half the bar diagonal; hole h is kept when
 * |u_h - u_d| <= sqrt(4R^2 + T^2)/(2 D_h) + r_d/D_d, with the hole centre at
 * the middle of its y span. */
static void pe_cull(const smg_o_geom* g, int d, const double p[3], cull_t* cl) {
    bar_t bd = get_bar(g, d);
    double dist = dist3(bd.c, p);
    double crit_self = sqrt(bd.w * bd.w + bd.h * bd.h + bd.t * bd.t) / dist / 2.0;
    double ud[3] = {(bd.c[0] - p[0]) / dist, (bd.c[1] - p[1]) / dist, (bd.c[2] - p[2]) / dist};
    cl->n_cryst = 0;
    cl->cryst = (int*)malloc(sizeof(int) * (g->n_det > 0 ? g->n_det : 1));
    for (int k = 0; k < g->n_det; ++k) {
        if (k == d) continue;
        bar_t bk = get_bar(g, k);
        double dk = dist3(p, bk.c);
        double crit = sqrt(bk.w * bk.w + bk.h * bk.h + bk.t * bk.t) / dk / 2.0;
        double uk[3] = {(bk.c[0] - p[0]) / dk, (bk.c[1] - p[1]) / dk, (bk.c[2] - p[2]) / dk};
        double s = sqrt((uk[2] - ud[2]) * (uk[2] - ud[2]) + (uk[1] - ud[1]) * (uk[1] - ud[1]) +
                        (uk[0] - ud[0]) * (uk[0] - ud[0]));
        if (s <= crit + crit_self) cl->cryst[cl->n_cryst++] = k;
    }
    int h = 0;
    for (int m = 0; m < g->n_layers; ++m) {
        cl->n_holes[m] = 0;
        cl->holes[m] = (int*)malloc(sizeof(int) * (g->holes_in_layer[m] > 0 ? g->holes_in_layer[m] : 1));
        for (int i = 0; i < g->holes_in_layer[m]; ++i, ++h) {
            const float* hp = g->coll + 100 + 9 * h;
            double xc = hp[0], y1 = (double)hp[1] + g->F, y2 = (double)hp[2] + g->F, zc = hp[3];
            double R = hp[4];
            double exr, ezr, eth;  /* an elliptical hole culls with its bounding radius */
            if (hole_ellipse(g, h, &exr, &ezr, &eth)) R = exr > ezr ? exr : ezr;
            double yc = (y1 + y2) / 2.0;
            double crit = sqrt(R * R * 4 + (y1 - y2) * (y1 - y2)) / 2.0;
            double dh = sqrt((p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc) +
                             (p[2] - zc) * (p[2] - zc));
            crit = crit / dh;
            double uh[3] = {(xc - p[0]) / dh, (yc - p[1]) / dh, (zc - p[2]) / dh};
            double s = sqrt((uh[2] - ud[2]) * (uh[2] - ud[2]) + (uh[1] - ud[1]) * (uh[1] - ud[1]) +
                            (uh[0] - ud[0]) * (uh[0] - ud[0]));
            if (s <= crit + crit_self) cl->holes[m][cl->n_holes[m]++] = h;
        }
    }
}

static void cull_free(const smg_o_geom* g, cull_t* cl) {
    free(cl->cryst);
    for (int m = 0; m < g->n_layers; ++m) free(cl->holes[m]);
}

void smg_o_pe_cull_counts(const smg_o_geom* g, int d, int v, int rot, int* n_cryst,
                          int* n_holes) {
    double p[3];
    voxel_pos(g, v, rot, p);
    cull_t cl;
    pe_cull(g, d, p, &cl);
    *n_cryst = cl.n_cryst;
    int nh = 0;
    for (int m = 0; m < g->n_layers; ++m) nh += cl.n_holes[m];
    *n_holes = nh;
    cull_free(g, &cl);
}

/* Attenuation of the ray p -> q up to the sub-cell [clo, chi] of bar d that
 * holds q (PE:1240-1346): plate gates, holes and plate material, the culled
 * bars, and the own-bar path l1 - l2.  lay_len and bar_len receive the layer
 * and bar path lengths (for the gradient), *l1o and *l2o the own-bar chords
 * from the bar's and the sub-cell's surface to q. */
static double pe_ray_att(const smg_o_geom* g, const bar_t* bdp, const cull_t* cl, const double p[3],
                         const double ps[3], const double q[3], const double qs[3],
                         const double clo[3], const double chi[3], double cosang, double* lay_len,
                         double* bar_len, double* l1o, double* l2o) {
    const bar_t bd = *bdp;
    const int Lc = g->n_layers;
    double blo[3] = {-bd.w / 2.0, -bd.t / 2.0, -bd.h / 2.0};
    double bhi[3] = {bd.w / 2.0, bd.t / 2.0, bd.h / 2.0};
    double att = 0.0;
    /* plate gate: a ray that does not cross a plate box is blocked (PE:1240-1254) */
    for (int m = 0; m < Lc; ++m) {
        const float* L = g->coll + (m + 1) * 10;
        double lo[3] = {-(double)L[1] / 2.0, -(double)L[2] / 2.0 + g->F + L[4], -(double)L[3] / 2.0};
        double hi[3] = {(double)L[1] / 2.0, (double)L[2] / 2.0 + g->F + L[4], (double)L[3] / 2.0};
        OPS(10); /* plate box */
        if (smg_o_box_ray(p, q, lo, hi) < (double)0.1f) att += 100000.0;
    }
    /* holes and plate material (PE:1257-1286) */
    for (int m = 0; m < Lc; ++m) {
        const float* L = g->coll + (m + 1) * 10;
        double in_coll = (double)L[2] / cosang;
        double sum = 0.0;
        OPS(5); /* plate path and its attenuation */
        for (int i = 0; i < cl->n_holes[m]; ++i) {
            const float* hp = g->coll + 100 + 9 * cl->holes[m][i];
            OPS(7);
            double exr, ezr, eth;
            double l = hole_ellipse(g, cl->holes[m][i], &exr, &ezr, &eth)
                           ? smg_o_ellcyl_ray(p, q, hp[0], (double)hp[1] + g->F,
                                              (double)hp[2] + g->F, hp[3], exr, ezr, eth)
                           : smg_o_cylinder_ray(p, q, hp[0], (double)hp[1] + g->F,
                                                (double)hp[2] + g->F, hp[3], hp[4]);
            sum += l;
            att += (double)hp[5] * l;
        }
        lay_len[m] = 0.0;
        if (in_coll - sum >= (double)0.00001f) {
            att += (double)L[5] * (in_coll - sum);
            lay_len[m] = in_coll - sum;
        }
    }
    /* other bars on the way (PE:1290-1325) */
    for (int i = 0; i < cl->n_cryst; ++i) {
        bar_t bk = get_bar(g, cl->cryst[i]);
        double pa[3], qa[3];
        to_bar_frame(&bk, p, pa);
        to_bar_frame(&bk, q, qa);
        double lo[3] = {-0.5 * bk.w, -0.5 * bk.t, -0.5 * bk.h};
        double hi[3] = {0.5 * bk.w, 0.5 * bk.t, 0.5 * bk.h};
        OPS(13); /* box, attenuation (the two frame transforms, 18, are
                  * counted by to_bar_frame) */
        double l = smg_o_box_ray(pa, qa, lo, hi);
        bar_len[i] = l;
        att += l * bk.mut;
    }
    /* own bar up to the sub-cell (PE:1329-1346) */
    OPS(21 + 2 + 20 + 3); /* cell box, exp, extension, absorption, sum */
    double l1 = smg_o_box_ray_inside(ps, qs, blo, bhi);
    double l2 = smg_o_box_ray_inside(ps, qs, clo, chi);
    *l1o = l1;
    *l2o = l2;
    return att + (l1 - l2) * bd.mut;
}

/* One element A[d][v] of the PE matrix (PE:1193-1362), summed over the
 * (sx, sy, sz) sub-cells of bar d.  When grad != NULL the derivatives of
 * wdv * A[d][v] with respect to the layer mu, the bars' mu_total and the
 * detection bar's mu_pe are added to grad (layout in the header). */
static double pe_element_core(const smg_o_geom* g, int d, int v, int rot, int sx, int sy,
                              int sz, double wdv, double* grad) {
    bar_t bd = get_bar(g, d);
    if (!(bd.mut > (double)0.01f)) return 0.0; /* PE:1116, PE:1194 */
    double p[3];
    voxel_pos(g, v, rot, p);
    /* the element set-up (voxel in the bar frame, culling scan) is not
     * counted: the op count is per sub-cell ray */
    const int counting = g_count_ops;
    g_count_ops = 0;
    double ps[3];
    to_bar_frame(&bd, p, ps);
    cull_t cl;
    pe_cull(g, d, p, &cl);
    g_count_ops = counting;
    const int Lc = g->n_layers;
    double cb = cos(bd.beta), sb = sin(bd.beta);
    double acc = 0.0;
    double* lay_len = (double*)malloc(sizeof(double) * (Lc > 0 ? Lc : 1));
    double* bar_len = (double*)malloc(sizeof(double) * (cl.n_cryst > 0 ? cl.n_cryst : 1));
    for (int c = 0; c < sz; ++c)
        for (int a = 0; a < sx; ++a)
            for (int bb = 0; bb < sy; ++bb) {
                double qs[3], clo[3], chi[3], q[3];
                cell_of(&bd, a, bb, c, sx, sy, sz, qs, clo, chi);
                from_bar_frame(&bd, qs, q);
                /* sub-cell point 24 (of which the turn into the world frame,
                 * 9, is counted by from_bar_frame), distance 10, solid angle 32 */
                OPS(57);
                double D[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
                double dsq = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
                double dd = sqrt(dsq);
                double cosang = D[1] / dd;
                /* solid angle: max of the three projected cell faces (PE:1223-1231) */
                double om_x = (bd.t * bd.h / sz / sy) / (4 * M_PI * dsq) * fabs(D[0] * cb - D[2] * sb) / dd;
                double om_z = (bd.w * bd.t / sx / sy) / (4 * M_PI * dsq) * fabs(D[0] * sb + D[2] * cb) / dd;
                double om_y = (bd.w * bd.h / sx / sz) / (4 * M_PI * dsq) * fabs(D[1]) / dd;
                double om = om_x >= om_y ? om_x : om_y;
                om = om >= om_z ? om : om_z;
                double l1, l2;
                double att = pe_ray_att(g, &bd, &cl, p, ps, q, qs, clo, chi, cosang, lay_len,
                                        bar_len, &l1, &l2);
                double eff = exp(-att);
                /* absorption along the ray extended 1000 mm through the cell (PE:1350-1357) */
                double ext[3];
                for (int k = 0; k < 3; ++k) ext[k] = qs[k] + 1000.0 * (qs[k] - ps[k]) / dd;
                double la = smg_o_box_ray(ps, ext, clo, chi);
                double ea = exp(-la * bd.mut);
                double absf = (1.0 - ea) * bd.mupe / bd.mut;
                double contrib = eff * om * absf;
                acc += contrib;
                if (grad) {
                    double cw = wdv * contrib;
                    for (int m = 0; m < Lc; ++m) grad[m] -= cw * lay_len[m];
                    for (int i = 0; i < cl.n_cryst; ++i) grad[Lc + cl.cryst[i]] -= cw * bar_len[i];
                    grad[Lc + d] -= cw * (l1 - l2);
                    double dabs_dmut = bd.mupe * (la * ea / bd.mut - (1.0 - ea) / (bd.mut * bd.mut));
                    grad[Lc + d] += wdv * eff * om * dabs_dmut;
                    grad[Lc + g->n_det + d] += wdv * eff * om * (1.0 - ea) / bd.mut;
                }
            }
    free(lay_len);
    free(bar_len);
    cull_free(g, &cl);
    return acc;
}

double smg_o_pe_element(const smg_o_geom* g, int d, int v, int rot, int sx, int sy, int sz) {
    return pe_element_core(g, d, v, rot, sx, sy, sz, 0.0, NULL);
}

double smg_o_pe_ops_per_ray(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* d,
                            const int* v, int n) {
    g_count_ops = 1;
    g_ops = 0.0;
    double rays = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!((double)g->det[12 * d[i] + 7] > (double)0.01f)) continue;
        pe_element_core(g, d[i], v[i], rot, sx, sy, sz, 0.0, NULL);
        rays += (double)sx * sy * sz;
    }
    g_count_ops = 0;
    return rays > 0 ? g_ops / rays : 0.0;
}

void smg_o_pe_rows(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* rows,
                   int nrows, double* out, int nthreads) {
    const int V = g->n_vox;
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) collapse(2)
    for (int i = 0; i < nrows; ++i)
        for (int v = 0; v < V; ++v)
            out[(size_t)i * V + v] = pe_element_core(g, rows[i], v, rot, sx, sy, sz, 0.0, NULL);
}

void smg_o_pe_pairs(const smg_o_geom* g, int rot, int sx, int sy, int sz, const int* d,
                    const int* v, int n, double* out, int nthreads) {
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < n; ++i) out[i] = pe_element_core(g, d[i], v[i], rot, sx, sy, sz, 0.0, NULL);
}

void smg_o_pe_grad(const smg_o_geom* g, int rot, int sx, int sy, int sz, const double* w,
                   double* grad, double* loss, int nthreads) {
    const int N = g->n_det, V = g->n_vox, P = g->n_layers + 2 * N;
    if (nthreads <= 0) nthreads = smg_o_max_threads();
    memset(grad, 0, sizeof(double) * P);
    double total = 0.0;
    /* one private gradient per detector row, summed in row order afterwards so
     * the result does not depend on the thread count */
    double* part = (double*)calloc((size_t)N * P, sizeof(double));
    double* lrow = (double*)calloc((size_t)N, sizeof(double));
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int d = 0; d < N; ++d) {
        double* gp = part + (size_t)d * P;
        for (int v = 0; v < V; ++v) {
            double wv = w[(size_t)d * V + v];
            if (wv == 0.0) continue; /* every term of this element is a multiple of wv */
            lrow[d] += wv * pe_element_core(g, d, v, rot, sx, sy, sz, wv, gp);
        }
    }
    for (int d = 0; d < N; ++d) {
        total += lrow[d];
        for (int i = 0; i < P; ++i) grad[i] += part[(size_t)d * P + i];
    }
    if (loss) *loss = total;
    free(part);
    free(lrow);
}

/* ======================================================================== */
/* Geometry bitmaps (SC:1040-1074, SC:1440-1540, SC:1939-2009)              */
/* ======================================================================== */

int64_t smg_o_bitmap_crystal_words(int n_det) {
    int64_t pairs = (int64_t)n_det * (n_det + 1) / 2;
    int64_t bits = pairs * n_det;
    return (bits + 31) / 32;
}

int64_t smg_o_bitmap_collimator_words(int n_holes_l0, int n_det) {
    int64_t bits = (int64_t)n_holes_l0 * n_det * n_det;
    return (bits + 31) / 32;
}

/* Bit of triple (i, j, k) for the unordered pair {i, j}: pair index
 * i*N - i(i+1)/2 + j with i <= j, bit pair*N + k, little bit order in
 * 32-bit words (SC:1040-1059). */
static int64_t crystal_bit_index(int n, int i, int j, int k) {
    if (i > j) {
        int t = i;
        i = j;
        j = t;
    }
    int64_t pair = (int64_t)i * n - ((int64_t)i * (i + 1)) / 2 + j;
    return pair * n + k;
}

int smg_o_bit_crystal(const uint32_t* words, int n_det, int i, int j, int k) {
    int64_t b = crystal_bit_index(n_det, i, j, k);
    return (words[b / 32] >> (b % 32)) & 1u;
}

int smg_o_bit_collimator(const uint32_t* words, int n_det, int i, int j, int k) {
    int64_t b = (int64_t)i * n_det * n_det + (int64_t)j * n_det + k;
    return (words[b / 32] >> (b % 32)) & 1u;
}

static void bar_center_raw(const smg_o_geom* g, int k, double c[3], double* r) {
    const float* d = g->det + 12 * k;
    c[0] = d[1];
    c[1] = d[2];
    c[2] = d[3];
    double w = d[4], h = d[6], t = d[5];
    *r = sqrt(w * w + h * h + t * t) / 2.0;
}

/* Bar k lies on the line between bars i < j if its bounding sphere, seen from
 * i, overlaps the direction to j by the angular test, or the same seen from j
 * (SC:1440-1540).  Positions omit F (a translation that cancels). */
int smg_o_crystal_pred(const smg_o_geom* g, int i, int j, int k) {
    double ci[3], ri, cj[3], rj, ck[3], rk;
    bar_center_raw(g, i, ci, &ri);
    bar_center_raw(g, j, cj, &rj);
    bar_center_raw(g, k, ck, &rk);
    double lij = dist3(ci, cj);
    double crit_i_j = ri / lij, crit_j_i = rj / lij;
    double uij[3] = {(ci[0] - cj[0]) / lij, (ci[1] - cj[1]) / lij, (ci[2] - cj[2]) / lij};
    double uji[3] = {(cj[0] - ci[0]) / lij, (cj[1] - ci[1]) / lij, (cj[2] - ci[2]) / lij};
    double lik = dist3(ck, ci), ljk = dist3(ck, cj);
    double uki[3] = {(ck[0] - ci[0]) / lik, (ck[1] - ci[1]) / lik, (ck[2] - ci[2]) / lik};
    double ukj[3] = {(ck[0] - cj[0]) / ljk, (ck[1] - cj[1]) / ljk, (ck[2] - cj[2]) / ljk};
    if (dist3(uki, uji) <= rk / lik + crit_j_i) return 1;
    return dist3(ukj, uij) <= rk / ljk + crit_i_j;
}

void smg_o_bitmap_crystal(const smg_o_geom* g, uint32_t* words, int nthreads) {
    const int N = g->n_det;
    int64_t nw = smg_o_bitmap_crystal_words(N);
    memset(words, 0, sizeof(uint32_t) * (size_t)nw);
    if (nthreads <= 0) nthreads = smg_o_max_threads();
    /* neighbouring pairs share a word, so bits are OR-ed atomically */
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (k == i || k == j || !smg_o_crystal_pred(g, i, j, k)) continue;
                int64_t b = crystal_bit_index(N, i, j, k);
#pragma omp atomic
                words[b / 32] |= (1u << (b % 32));
            }
}

/* Bar k lies on the ray from layer-0 hole i toward bar j (SC:1939-2009):
 * angular test at the hole centre (x, (y1+y2)/2, z), F omitted. */
int smg_o_coll_pred(const smg_o_geom* g, int i, int j, int k) {
    const float* hp = g->coll + 100 + 9 * i;
    double ci[3] = {hp[0], ((double)hp[1] + (double)hp[2]) / 2.0, hp[3]};
    double cj[3], rj, ck[3], rk;
    bar_center_raw(g, j, cj, &rj);
    bar_center_raw(g, k, ck, &rk);
    double lij = dist3(ci, cj);
    double crit_j_i = rj / lij;
    double uji[3] = {(cj[0] - ci[0]) / lij, (cj[1] - ci[1]) / lij, (cj[2] - ci[2]) / lij};
    double lik = dist3(ck, ci);
    double uki[3] = {(ck[0] - ci[0]) / lik, (ck[1] - ci[1]) / lik, (ck[2] - ci[2]) / lik};
    return dist3(uki, uji) <= rk / lik + crit_j_i;
}

void smg_o_bitmap_collimator(const smg_o_geom* g, uint32_t* words, int nthreads) {
    const int N = g->n_det, H = g->n_holes_l0;
    int64_t nw = smg_o_bitmap_collimator_words(H, N);
    memset(words, 0, sizeof(uint32_t) * (size_t)nw);
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) collapse(2)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (!smg_o_coll_pred(g, i, j, k)) continue;
                int64_t b = (int64_t)i * N * N + (int64_t)j * N + k;
#pragma omp atomic
                words[b / 32] |= (1u << (b % 32));
            }
}

/* ======================================================================== */
/* mu(E) tables (an extension with no reference counterpart)               */
/* ======================================================================== */

/* Tables on increasing node energies (keV), one per process; interpolation
 * is linear in (log E, log mu) between nodes and continues the end segments
 * beyond them.  n = 0 returns to the reference's power laws. */
int smg_o_set_mu_tables(smg_o_geom* g, int n, const double* energy, const double* crystal_pe,
                        const double* crystal_c, const double* coll_pe, const double* coll_c) {
    if (n == 0) {
        g->n_tab = 0;
        return 0;
    }
    if (n < 2 || n > SMG_O_MAX_NODES) return -1;
    const double* v[4] = {crystal_pe, crystal_c, coll_pe, coll_c};
    for (int i = 0; i < n; ++i) {
        if (!(energy[i] > 0.0) || (i > 0 && !(energy[i] > energy[i - 1]))) return -1;
        g->tab_le[i] = log(energy[i]);
        for (int t = 0; t < 4; ++t) {
            if (!(v[t][i] > 0.0)) return -1;
            g->tab_lmu[t][i] = log(v[t][i]);
        }
    }
    g->n_tab = n;
    return 0;
}

static double tab_lmu_at(const smg_o_geom* g, int which, double le) {
    int lo = 0, hi = g->n_tab - 1;
    while (hi - lo > 1) { /* the last node not above le, within [0, n - 2] */
        int mid = (lo + hi) / 2;
        if (g->tab_le[mid] <= le) lo = mid;
        else hi = mid;
    }
    double f = (le - g->tab_le[lo]) / (g->tab_le[lo + 1] - g->tab_le[lo]);
    return g->tab_lmu[which][lo] + f * (g->tab_lmu[which][lo + 1] - g->tab_lmu[which][lo]);
}

/* mu(e) / mu(e0) of process `which` (0 crystal pe, 1 crystal compton,
 * 2 collimator pe, 3 collimator compton): from the tables when set, else the
 * reference's scaling (e0/e)^3 for photoelectric and e0/e for Compton
 * (SC:1371-1375). */
double smg_o_mu_ratio(const smg_o_geom* g, int which, double e, double e0) {
    if (g->n_tab == 0) {
        OPS((which & 1) ? 1 : 3);
        double r = e0 / e;
        return (which & 1) ? r : r * r * r;
    }
    OPS(18); /* two logs, two interpolations, difference, exp */
    return exp(tab_lmu_at(g, which, log(e)) - tab_lmu_at(g, which, log(e0)));
}

/* ======================================================================== */
/* Primary Compton scatter (SC:1077-1437, SC:1544-1936, SCC:1207-1215)      */
/* ======================================================================== */

/* ---- geometric chords for the grid-scatterer extension ----------------- */
/* These have no reference counterpart: plain intersection lengths of a
 * segment with a box and with a finite cylinder along y, without the
 * reference chord functions' thresholds. */

/* Length of the segment a -> b inside the box [lo, hi]. */
double smg_o_seg_box(const double a[3], const double b[3], const double lo[3], const double hi[3]) {
    OPS(16); /* per axis: difference, two offsets, two quotients; the length */
    double t0 = 0.0, t1 = 1.0;
    for (int k = 0; k < 3; ++k) {
        double D = b[k] - a[k];
        if (D == 0.0) {
            if (a[k] < lo[k] || a[k] > hi[k]) return 0.0;
            continue;
        }
        double u = (lo[k] - a[k]) / D, w = (hi[k] - a[k]) / D;
        if (u > w) { double x = u; u = w; w = x; }
        if (u > t0) t0 = u;
        if (w < t1) t1 = w;
    }
    return t1 > t0 ? (t1 - t0) * dist3(a, b) : 0.0;
}

/* Length of the segment a -> b inside the cylinder of radius r about the
 * line x = xc, z = zc, for y1 <= y <= y2. */
double smg_o_seg_cyl(const double a[3], const double b[3], double xc, double y1, double y2,
                     double zc, double r) {
    OPS(35); /* direction, y slab, offsets, quadratic, sqrt, roots, length */
    double D[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double t0 = 0.0, t1 = 1.0;
    if (D[1] == 0.0) {
        if (a[1] < y1 || a[1] > y2) return 0.0;
    } else {
        double u = (y1 - a[1]) / D[1], w = (y2 - a[1]) / D[1];
        if (u > w) { double x = u; u = w; w = x; }
        if (u > t0) t0 = u;
        if (w < t1) t1 = w;
    }
    double x0 = a[0] - xc, z0 = a[2] - zc;
    double A = D[0] * D[0] + D[2] * D[2], B = x0 * D[0] + z0 * D[2], Cq = x0 * x0 + z0 * z0 - r * r;
    if (A == 0.0) {
        if (Cq > 0.0) return 0.0;
    } else {
        double disc = B * B - A * Cq;
        if (disc <= 0.0) return 0.0;
        double sq = sqrt(disc);
        double u = (-B - sq) / A, w = (-B + sq) / A;
        if (u > t0) t0 = u;
        if (w < t1) t1 = w;
    }
    return t1 > t0 ? (t1 - t0) * dist3(a, b) : 0.0;
}

/* Length of the segment a -> b inside the elliptic cylinder of hole shape
 * (x_rho, z_rho, theta) about the line x = xc, z = zc, for y1 <= y <= y2: the
 * point's offsets from the axis turned as in smg_o_ellcyl_ray,
 * x' = dx cos(theta) + dz sin(theta), z' = -dx sin(theta) + dz cos(theta),
 * lie in x'^2 / x_rho^2 + z'^2 / z_rho^2 <= 1.  Along the segment x', z' are
 * linear in the parameter t, so the condition is the quadratic
 * q2 t^2 + 2 q1 t + q0 <= 0 solved here directly. */
double smg_o_seg_ell(const double a[3], const double b[3], double xc, double y1, double y2,
                     double zc, double x_rho, double z_rho, double theta) {
    OPS(50); /* turn of the start and direction, y slab, quadratic, sqrt, roots, length */
    double cs = cos(theta), sn = sin(theta);
    double D[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double t0 = 0.0, t1 = 1.0;
    if (D[1] == 0.0) {
        if (a[1] < y1 || a[1] > y2) return 0.0;
    } else {
        double u = (y1 - a[1]) / D[1], w = (y2 - a[1]) / D[1];
        if (u > w) { double x = u; u = w; w = x; }
        if (u > t0) t0 = u;
        if (w < t1) t1 = w;
    }
    double dx = a[0] - xc, dz = a[2] - zc;
    double px = dx * cs + dz * sn, pz = -dx * sn + dz * cs;       /* start, turned */
    double vx = D[0] * cs + D[2] * sn, vz = -D[0] * sn + D[2] * cs; /* direction, turned */
    double ix = 1.0 / (x_rho * x_rho), iz = 1.0 / (z_rho * z_rho);
    double q2 = vx * vx * ix + vz * vz * iz;
    double q1 = px * vx * ix + pz * vz * iz;
    double q0 = px * px * ix + pz * pz * iz - 1.0;
    if (q2 == 0.0) {
        if (q0 > 0.0) return 0.0;
    } else {
        double disc = q1 * q1 - q2 * q0;
        if (disc <= 0.0) return 0.0;
        double sq = sqrt(disc);
        double u = (-q1 - sq) / q2, w = (-q1 + sq) / q2;
        if (u > t0) t0 = u;
        if (w < t1) t1 = w;
    }
    return t1 > t0 ? (t1 - t0) * dist3(a, b) : 0.0;
}

/* Whether grid point (x, y, z) lies inside hole h (strictly inside its
 * circle or ellipse, and within its y span). */
static int in_hole(const smg_o_geom* g, int h, double x, double y, double z) {
    const float* hp = g->coll + 100 + 9 * h;
    double dx = x - hp[0], dz = z - hp[3];
    double exr, ezr, eth;
    int in;
    if (hole_ellipse(g, h, &exr, &ezr, &eth)) {
        double cs = cos(eth), sn = sin(eth);
        double u = dx * cs + dz * sn, w = -dx * sn + dz * cs;
        in = u * u / (exr * exr) + w * w / (ezr * ezr) < 1.0;
    } else {
        in = dx * dx + dz * dz < (double)hp[4] * hp[4];
    }
    return in && y >= (double)hp[1] + g->F && y <= (double)hp[2] + g->F;
}

/* Collimator material along the segment a -> b at E0: for each layer whose
 * plate box the segment meets, its plate chord less the chords of its holes
 * (at least 0) with the layer's mu, plus the holes' own mu.  att_t = sum mu_total l; s_pe, s_c the sums with
 * the photoelectric and Compton coefficients (for the scaling to E'). */
static void coll_material(const smg_o_geom* g, const double a[3], const double b[3],
                          double* att_t, double* s_pe, double* s_c) {
    *att_t = *s_pe = *s_c = 0.0;
    int h0 = 0;
    for (int m = 0; m < g->n_layers; ++m) {
        const float* L = g->coll + (m + 1) * 10;
        double lo[3] = {-(double)L[1] / 2.0, -(double)L[2] / 2.0 + g->F + L[4], -(double)L[3] / 2.0};
        double hi[3] = {(double)L[1] / 2.0, (double)L[2] / 2.0 + g->F + L[4], (double)L[3] / 2.0};
        double c = smg_o_seg_box(a, b, lo, hi);
        if (!(c > 0.0)) { /* the layer's holes lie inside its plate box */
            h0 += g->holes_in_layer[m];
            continue;
        }
        double holes = 0.0;
        for (int i = 0; i < g->holes_in_layer[m]; ++i) {
            const float* hp = g->coll + 100 + 9 * (h0 + i);
            double exr, ezr, eth;
            double l = hole_ellipse(g, h0 + i, &exr, &ezr, &eth)
                           ? smg_o_seg_ell(a, b, hp[0], (double)hp[1] + g->F, (double)hp[2] + g->F, hp[3],
                                           exr, ezr, eth)
                           : smg_o_seg_cyl(a, b, hp[0], (double)hp[1] + g->F, (double)hp[2] + g->F, hp[3],
                                           hp[4]);
            holes += l;
            *att_t += (double)hp[5] * l;
            *s_pe += (double)hp[6] * l;
            *s_c += (double)hp[7] * l;
        }
        h0 += g->holes_in_layer[m];
        double wl = c - holes > 0.0 ? c - holes : 0.0;
        *att_t += (double)L[5] * wl;
        *s_pe += (double)L[6] * wl;
        *s_c += (double)L[7] * wl;
    }
}

/* Scatter points on a grid over the plates of the layers in layer_mask (bit
 * m: layer m): per layer, n_k = max(1, round(extent / pitch)) cells per axis,
 * points at the cell centres, z outermost and x innermost, without the points
 * inside a hole of that layer; layers in increasing order.  Writes up to max
 * points (x, y, z, w) when xyzw is given, w = cell volume x the layer's
 * Compton mu; returns the number of points. */
int64_t smg_o_coll_grid_points(const smg_o_geom* g, double pitch, unsigned layer_mask, double* xyzw,
                               int64_t max) {
    if (!(pitch > 0.0)) return 0;
    int64_t cnt = 0;
    int first = 0;
    for (int m = 0; m < g->n_layers; first += g->holes_in_layer[m], ++m) {
        if (!((layer_mask >> m) & 1u)) continue;
        const float* L = g->coll + (m + 1) * 10;
        double lo[3] = {-(double)L[1] / 2.0, -(double)L[2] / 2.0 + g->F + L[4], -(double)L[3] / 2.0};
        double ext[3] = {L[1], L[2], L[3]};
        int n[3];
        double pk[3];
        for (int k = 0; k < 3; ++k) {
            n[k] = (int)floor(ext[k] / pitch + 0.5);
            if (n[k] < 1) n[k] = 1;
            pk[k] = ext[k] / n[k];
        }
        const double w = pk[0] * pk[1] * pk[2] * (double)L[7];
        for (int kz = 0; kz < n[2]; ++kz)
            for (int jy = 0; jy < n[1]; ++jy)
                for (int ix = 0; ix < n[0]; ++ix) {
                    double x = lo[0] + (ix + 0.5) * pk[0], y = lo[1] + (jy + 0.5) * pk[1],
                           z = lo[2] + (kz + 0.5) * pk[2];
                    int inside = 0;
                    for (int h = first; h < first + g->holes_in_layer[m] && !inside; ++h)
                        inside = in_hole(g, h, x, y, z);
                    if (inside) continue;
                    if (xyzw && cnt < max) {
                        xyzw[4 * cnt] = x;
                        xyzw[4 * cnt + 1] = y;
                        xyzw[4 * cnt + 2] = z;
                        xyzw[4 * cnt + 3] = w;
                    }
                    ++cnt;
                }
    }
    return cnt;
}

/* First-leg weight of grid scatter point s (weight w = dv mu_c of its layer)
 * for a source at p: the chance of a Compton interaction in the cell around
 * s, w / (4 pi |s - p|^2) exp(-collimator material along p -> s). */
double smg_o_coll_grid_weight(const smg_o_geom* g, const double p[3], const double s[3], double w) {
    double att_t, s_pe, s_c;
    coll_material(g, p, s, &att_t, &s_pe, &s_c);
    double L = dist3(p, s);
    return w / (4.0 * M_PI * L * L) * exp(-att_t);
}

typedef struct {
    double e0, lo, hi, norm;
    int compton;
} phys_t;

static phys_t phys_for(const smg_o_geom* g, const bar_t* bd) {
    phys_t ph;
    ph.compton = (int)floor((double)g->phys[0] + 0.5);
    int same = (int)floor((double)g->phys[4] + 0.5);
    ph.e0 = g->phys[7];
    if (same > 0) {
        ph.lo = g->phys[5];
        ph.hi = g->phys[6];
    } else { /* per-bar window E0 (1 -+ res/2) (SC:1154-1163) */
        ph.lo = (1 - bd->res / 2.0) * ph.e0;
        ph.hi = (1 + bd->res / 2.0) * ph.e0;
    }
    ph.norm = smg_o_kn_sum(ph.e0, 0.0, KN_PI, KN_STEP); /* SC:1169 */
    return ph;
}

/* Scatter from point qsw (world) toward the sub-cells of bar d, for voxel
 * position p, weighted by wgt (SC:1226-1428).  kind 0: scatterer is bar sidx
 * (crystal bitmap); kind 1: scatterer is layer-0 hole sidx (collimator
 * bitmap). */
static double scatter_body(const smg_o_geom* g, const bar_t* bd, int d, const double p[3],
                           const double qsw[3], double wgt, const uint32_t* bitmap, int kind,
                           int sidx, int sx, int sy, int sz, const phys_t* ph) {
    const int N = g->n_det;
    double ps[3], qss[3];
    to_bar_frame(bd, p, ps);
    to_bar_frame(bd, qsw, qss);
    double blo[3] = {-bd->w / 2.0, -bd->t / 2.0, -bd->h / 2.0};
    double bhi[3] = {bd->w / 2.0, bd->t / 2.0, bd->h / 2.0};
    double acc = 0.0;
    for (int c = 0; c < sz; ++c)
        for (int a = 0; a < sx; ++a)
            for (int bb = 0; bb < sy; ++bb) {
                double qs[3], clo[3], chi[3], q[3];
                if (g_count_ops) g_terms += 1.0;
                OPS(30); /* cell centre and box */
                cell_of(bd, a, bb, c, sx, sy, sz, qs, clo, chi);
                from_bar_frame(bd, qs, q);
                double theta = smg_o_cone_angle(p, qsw, q);
                double ep = smg_o_scatter_energy(theta, ph->e0);
                OPS(3 + 8); /* resolution at E', the two window tests */
                double resp = bd->res * sqrt(ep / ph->e0);
                /* 2-sigma energy-window early out (SC:1263-1269) */
                if (ep * (1 + 2 * resp / FWHM_SIGMA) <= ph->lo || ep * (1 - 2 * resp / FWHM_SIGMA) >= ph->hi)
                    continue;
                double pe_win = smg_o_gauss_window(ep, resp, ph->lo, ph->hi);
                double L = dist3(qsw, q);
                OPS(13 + 3); /* cell radius, the 2 pi test */
                double rdet = sqrt(bd->w * bd->w / sx / sx + bd->h * bd->h / sz / sz +
                                   bd->t * bd->t / sy / sy) / 2.0;
                double phi;
                if (L * sin(theta) * 2.0 <= rdet) {
                    phi = 2.0 * M_PI;
                } else {
                    OPS(6);
                    double r = rdet / 2.0 / L / sin(theta);
                    phi = 4.0 * asin(r <= 1.0 ? r : 1.0);
                }
                /* polar range: cone angles to the 8 cell corners (SC:1290-1335) */
                double dis = dist3(qss, ps);
                double ext[3];
                OPS(12);
                for (int k = 0; k < 3; ++k) ext[k] = qss[k] + 1000.0 * (qss[k] - ps[k]) / dis;
                double lthru = smg_o_box_ray(ps, ext, clo, chi);
                double tmin = 0.0, tmax = 0.0;
                for (int corner = 0; corner < 8; ++corner) {
                    double cc[3] = {(corner & 1) ? chi[0] : clo[0], (corner & 2) ? chi[1] : clo[1],
                                    (corner & 4) ? chi[2] : clo[2]};
                    double th = smg_o_cone_angle(ps, qss, cc);
                    if (corner == 0 || th > tmax) tmax = th;
                    if (corner == 0 || th < tmin) tmin = th;
                }
                if (lthru > (double)0.001f) tmin = 0.0;
                OPS(1);
                double ratio = smg_o_kn_sum(ph->e0, tmin, tmax, KN_STEP) / ph->norm;
                /* second leg: bars flagged in the bitmap, mu scaled to E' (SC:1343-1401) */
                double e3, e1; /* mu(E')/mu(E0): photoelectric, Compton (crystal) */
                if (g->n_tab == 0) {
                    OPS(6);
                    e3 = ph->e0 * ph->e0 * ph->e0 / ep / ep / ep;
                    e1 = ph->e0 / ep;
                } else {
                    e3 = smg_o_mu_ratio(g, 0, ep, ph->e0);
                    e1 = smg_o_mu_ratio(g, 1, ep, ph->e0);
                }
                double att = 0.0;
                for (int k = 0; k < N; ++k) {
                    if (k == d) continue;
                    int bit;
                    if (kind == 0 || kind == 3) {
                        if (k == sidx) continue;
                        bit = smg_o_bit_crystal(bitmap, N, sidx, d, k);
                    } else if (kind == 1) {
                        bit = smg_o_bit_collimator(bitmap, N, sidx, d, k);
                    } else {
                        bit = 1; /* grid scatterers: every other bar */
                    }
                    if (!bit) continue;
                    bar_t bk = get_bar(g, k);
                    OPS(3 + 2); /* mu at E', its product with the chord and the sum */
                    double mut = bk.mupe * e3 + bk.muc * e1;
                    double sa[3], qa[3];
                    to_bar_frame(&bk, qsw, sa);
                    to_bar_frame(&bk, q, qa);
                    double lo[3] = {-0.5 * bk.w, -0.5 * bk.t, -0.5 * bk.h};
                    double hi[3] = {0.5 * bk.w, 0.5 * bk.t, 0.5 * bk.h};
                    att += smg_o_box_ray(sa, qa, lo, hi) * mut;
                }
                if (kind == 3) { /* grid point inside bar sidx: its way out of that bar */
                    bar_t bs = get_bar(g, sidx);
                    double sa[3], qa[3];
                    to_bar_frame(&bs, qsw, sa);
                    to_bar_frame(&bs, q, qa);
                    double lo[3] = {-0.5 * bs.w, -0.5 * bs.t, -0.5 * bs.h};
                    double hi[3] = {0.5 * bs.w, 0.5 * bs.t, 0.5 * bs.h};
                    OPS(5);
                    att += smg_o_seg_box(sa, qa, lo, hi) * (bs.mupe * e3 + bs.muc * e1);
                }
                if (kind == 2) { /* collimator material on the second leg, at E' */
                    double att_t, s_pe, s_c;
                    coll_material(g, qsw, q, &att_t, &s_pe, &s_c);
                    att += s_pe * smg_o_mu_ratio(g, 2, ep, ph->e0) + s_c * smg_o_mu_ratio(g, 3, ep, ph->e0);
                }
                double l1 = smg_o_box_ray_inside(qss, qs, blo, bhi);
                double l2 = smg_o_box_ray_inside(qss, qs, clo, chi);
                OPS(4 + 3 + 12 + 14); /* own-bar mu at E', its path, extension, the term */
                double mupe_d = bd->mupe * e3, mut_d = mupe_d + bd->muc * e1;
                att += (l1 - l2) * mut_d;
                double ext2[3];
                for (int k = 0; k < 3; ++k) ext2[k] = qs[k] + 1000.0 * (qs[k] - qss[k]) / L;
                double la = smg_o_box_ray(qss, ext2, clo, chi);
                acc += wgt * phi / 2.0 / M_PI * ratio * pe_win * exp(-att) *
                       (1.0 - exp(-la * mut_d)) * mupe_d / mut_d;
            }
    return acc;
}

/* Element of the inter-crystal scatter matrix: sum over scatter bars s != d
 * at their centres, weighted by PE[s][v] * mu_c(d)/mu_pe(d) (SC:1077-1437). */
double smg_o_crystal_scatter_element(const smg_o_geom* g, const float* pe,
                                     const uint32_t* bitmap, int d, int v, int rot, int sx,
                                     int sy, int sz) {
    bar_t bd = get_bar(g, d);
    phys_t ph = phys_for(g, &bd);
    if (ph.compton != 1 || !(bd.mut > (double)0.01f)) return 0.0;
    double p[3];
    voxel_pos(g, v, rot, p);
    double acc = 0.0;
    for (int s = 0; s < g->n_det; ++s) {
        if (s == d) continue;
        bar_t bs = get_bar(g, s);
        double wgt = (double)pe[(size_t)s * g->n_vox + v] * bd.muc / bd.mupe;
        acc += scatter_body(g, &bd, d, p, bs.c, wgt, bitmap, 0, s, sx, sy, sz, &ph);
    }
    return acc;
}

/* Element of the collimator scatter matrix: one scatterer at the centre of
 * each layer-0 hole, weighted by a virtual-disc solid angle times
 * mu_c/mu_total of layer 0 (SC:1702-1716; standalone variant SCC:1207-1215). */
double smg_o_collimator_scatter_element(const smg_o_geom* g, const uint32_t* bitmap, int d,
                                        int v, int rot, int variant, int sx, int sy, int sz) {
    bar_t bd = get_bar(g, d);
    phys_t ph = phys_for(g, &bd);
    if (ph.compton != 1 || !(bd.mut > (double)0.01f)) return 0.0;
    double p[3];
    voxel_pos(g, v, rot, p);
    double acc = 0.0;
    for (int h = 0; h < g->n_holes_l0; ++h) {
        const float* hp = g->coll + 100 + 9 * h;
        double qsw[3] = {hp[0], ((double)hp[1] + (double)hp[2]) / 2.0 + g->F, hp[3]};
        double L = dist3(p, qsw);
        double T = (double)hp[2] - (double)hp[1];
        double om;
        if (variant == 0) {
            double r = (double)0.1f;
            om = (r * r * M_PI + T * r) / (4.0 * M_PI * L * L) * fabs(qsw[1] - p[1]) / L;
        } else {
            double r = (double)0.2f;
            om = (r * r + T * r) / (4.0 * L * L);
        }
        double wgt = om / (double)g->coll[15] * (double)g->coll[17];
        acc += scatter_body(g, &bd, d, p, qsw, wgt, bitmap, 1, h, sx, sy, sz, &ph);
    }
    return acc;
}

/* Element of the collimator scatter matrix with scatter points on a grid
 * over the tungsten (an extension with no reference counterpart): first-leg
 * weight smg_o_coll_grid_weight, the second leg through every other bar and
 * through the collimator material at E', the rest as the reference's
 * scatter element (SC:1226-1428).  pts: (x, y, z, w) per point. */
double smg_o_coll_grid_scatter_element(const smg_o_geom* g, const double* pts, int64_t n_pts, int d,
                                       int v, int rot, int sx, int sy, int sz) {
    bar_t bd = get_bar(g, d);
    phys_t ph = phys_for(g, &bd);
    if (ph.compton != 1 || !(bd.mut > (double)0.01f)) return 0.0;
    double p[3];
    voxel_pos(g, v, rot, p);
    double acc = 0.0;
    for (int64_t i = 0; i < n_pts; ++i) {
        const double* s = pts + 4 * i;
        double w = smg_o_coll_grid_weight(g, p, s, s[3]);
        acc += scatter_body(g, &bd, d, p, s, w, NULL, 2, -1, sx, sy, sz, &ph);
    }
    return acc;
}

void smg_o_coll_grid_scatter_pairs(const smg_o_geom* g, double pitch, unsigned layer_mask, int rot,
                                   int sx, int sy, int sz, const int* d, const int* v, int n,
                                   double* out, int nthreads) {
    int64_t np = smg_o_coll_grid_points(g, pitch, layer_mask, NULL, 0);
    double* pts = (double*)malloc(sizeof(double) * 4 * (size_t)(np > 0 ? np : 1));
    smg_o_coll_grid_points(g, pitch, layer_mask, pts, np);
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < n; ++i)
        out[i] = smg_o_coll_grid_scatter_element(g, pts, np, d[i], v[i], rot, sx, sy, sz);
    free(pts);
}

/* First-leg weight of the centre s of grid cell `cell` of bar b (cells of
 * a gx x gy x gz split in the PE sub-cell order) for voxel v: the chance of
 * a Compton interaction in that cell, dv mu_c,b / (4 pi |s - p|^2) exp(-att),
 * att being the PE ray's attenuation to s (PE:1240-1346 with the sub-cell's
 * own chord added back).  An extension with no reference counterpart. */
static double grid_weight_culled(const smg_o_geom* g, const bar_t* bbp, const cull_t* clp,
                                 const double p[3], const double ps[3], int gx, int gy, int gz,
                                 int cell);

double smg_o_crystal_grid_weight(const smg_o_geom* g, int b, int v, int rot, int gx, int gy,
                                 int gz, int cell) {
    bar_t bb = get_bar(g, b);
    if (!(bb.mut > (double)0.01f)) return 0.0;
    double p[3], ps[3];
    voxel_pos(g, v, rot, p);
    to_bar_frame(&bb, p, ps);
    cull_t cl;
    pe_cull(g, b, p, &cl);
    double w = grid_weight_culled(g, &bb, &cl, p, ps, gx, gy, gz, cell);
    cull_free(g, &cl);
    return w;
}

/* The weight above with bar b's culling lists for the voxel already formed
 * (they depend on the bar and the voxel only, not on the cell). */
static double grid_weight_culled(const smg_o_geom* g, const bar_t* bbp, const cull_t* clp,
                                 const double p[3], const double ps[3], int gx, int gy, int gz,
                                 int cell) {
    const bar_t bb = *bbp;
    const cull_t cl = *clp;
    const int by = cell % gy, a = (cell / gy) % gx, c = cell / (gy * gx);
    double qs[3], clo[3], chi[3], q[3];
    cell_of(&bb, a, by, c, gx, gy, gz, qs, clo, chi);
    from_bar_frame(&bb, qs, q);
    double D[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double dsq = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    double cosang = D[1] / sqrt(dsq);
    double* lay_len = (double*)malloc(sizeof(double) * (g->n_layers > 0 ? g->n_layers : 1));
    double* bar_len = (double*)malloc(sizeof(double) * (cl.n_cryst > 0 ? cl.n_cryst : 1));
    double l1, l2;
    double att = pe_ray_att(g, &bb, &cl, p, ps, q, qs, clo, chi, cosang, lay_len, bar_len, &l1, &l2);
    free(lay_len);
    free(bar_len);
    double dv = bb.w * bb.t * bb.h / ((double)gx * gy * gz);
    return exp(-(att + l2 * bb.mut)) * dv * bb.muc / (4.0 * M_PI * dsq);
}

/* Element of the crystal scatter matrix with scatter points at the grid
 * cell centres of every other bar (an extension with no reference
 * counterpart): first-leg weight smg_o_crystal_grid_weight, the second leg
 * through the bars flagged in the crystal bitmap for the bar pair and out
 * of the scatter bar itself, the rest as the reference's crystal scatter
 * element (SC:1077-1437). */
double smg_o_crystal_grid_scatter_element(const smg_o_geom* g, const uint32_t* bitmap, int d,
                                          int v, int rot, int gx, int gy, int gz, int sx, int sy,
                                          int sz) {
    bar_t bd = get_bar(g, d);
    phys_t ph = phys_for(g, &bd);
    if (ph.compton != 1 || !(bd.mut > (double)0.01f)) return 0.0;
    double p[3];
    voxel_pos(g, v, rot, p);
    double acc = 0.0;
    const int nc = gx * gy * gz;
    for (int b = 0; b < g->n_det; ++b) {
        if (b == d) continue;
        bar_t bb = get_bar(g, b);
        const int live = bb.mut > (double)0.01f;
        double pb[3];
        cull_t cl;
        if (live) {  /* bar b's culling for this voxel, shared by its cells */
            to_bar_frame(&bb, p, pb);
            pe_cull(g, b, p, &cl);
        }
        for (int cell = 0; cell < nc; ++cell) {
            const int by = cell % gy, a = (cell / gy) % gx, c = cell / (gy * gx);
            double qs[3], clo[3], chi[3], s[3];
            cell_of(&bb, a, by, c, gx, gy, gz, qs, clo, chi);
            from_bar_frame(&bb, qs, s);
            double w = live ? grid_weight_culled(g, &bb, &cl, p, pb, gx, gy, gz, cell) : 0.0;
            acc += scatter_body(g, &bd, d, p, s, w, bitmap, 3, b, sx, sy, sz, &ph);
        }
        if (live) cull_free(g, &cl);
    }
    return acc;
}

void smg_o_crystal_grid_scatter_pairs(const smg_o_geom* g, const uint32_t* bitmap, int rot,
                                      int gx, int gy, int gz, int sx, int sy, int sz,
                                      const int* d, const int* v, int n, double* out,
                                      int nthreads) {
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < n; ++i)
        out[i] = smg_o_crystal_grid_scatter_element(g, bitmap, d[i], v[i], rot, gx, gy, gz, sx,
                                                    sy, sz);
}

void smg_o_crystal_scatter_pairs(const smg_o_geom* g, const float* pe, const uint32_t* bitmap,
                                 int rot, int sx, int sy, int sz, const int* d, const int* v,
                                 int n, double* out, int nthreads) {
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < n; ++i)
        out[i] = smg_o_crystal_scatter_element(g, pe, bitmap, d[i], v[i], rot, sx, sy, sz);
}

void smg_o_collimator_scatter_pairs(const smg_o_geom* g, const uint32_t* bitmap, int rot,
                                    int variant, int sx, int sy, int sz, const int* d,
                                    const int* v, int n, double* out, int nthreads) {
    if (nthreads <= 0) nthreads = smg_o_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < n; ++i)
        out[i] = smg_o_collimator_scatter_element(g, bitmap, d[i], v[i], rot, variant, sx, sy, sz);
}

/* Algorithmic operations per scatter term (one (detector row, voxel,
 * scatter point, detection cell) term of the scatter element, SC:1226-1428):
 * the element body of the reference formula on the branches it takes, for
 * the rows d[i] and voxels v[i], counted on the calling thread, with unit
 * weights (the first-leg weights of the grid scatterers, their own kernels'
 * work, are not counted).  kind: 0 crystal scatter (other bar centres, crystal
 * bitmap), 1 collimator scatter (layer-0 hole centres, collimator bitmap),
 * 2 collimator grid (points of a pitch grid over the layers in layer_mask),
 * 3 crystal grid (centres of gx x gy x gz cells of the other bars, crystal
 * bitmap).  *terms receives the number of terms. */
double smg_o_scatter_ops_per_term(const smg_o_geom* g, const uint32_t* bitmap, int kind,
                                  double pitch, unsigned layer_mask, int gx, int gy, int gz, int rot,
                                  int sx, int sy, int sz, const int* d, const int* v, int n,
                                  double* terms) {
    double* pts = NULL;
    int64_t np = 0;
    if (kind == 2) {
        np = smg_o_coll_grid_points(g, pitch, layer_mask, NULL, 0);
        pts = (double*)malloc(sizeof(double) * 4 * (size_t)(np > 0 ? np : 1));
        smg_o_coll_grid_points(g, pitch, layer_mask, pts, np);
    }
    g_count_ops = 1;
    g_ops = 0.0;
    g_terms = 0.0;
    for (int i = 0; i < n; ++i) {
        bar_t bd = get_bar(g, d[i]);
        phys_t ph = phys_for(g, &bd);
        if (ph.compton != 1 || !(bd.mut > (double)0.01f)) continue;
        double p[3];
        voxel_pos(g, v[i], rot, p);
        if (kind == 0 || kind == 3) {
            const int nc = kind == 0 ? 1 : gx * gy * gz;
            for (int b = 0; b < g->n_det; ++b) {
                if (b == d[i]) continue;
                bar_t bb = get_bar(g, b);
                for (int cell = 0; cell < nc; ++cell) {
                    double s3[3];
                    if (kind == 0) {
                        s3[0] = bb.c[0]; s3[1] = bb.c[1]; s3[2] = bb.c[2];
                    } else {
                        const int by = cell % gy, a = (cell / gy) % gx, c = cell / (gy * gx);
                        double qs[3], clo[3], chi[3];
                        cell_of(&bb, a, by, c, gx, gy, gz, qs, clo, chi);
                        from_bar_frame(&bb, qs, s3);
                    }
                    scatter_body(g, &bd, d[i], p, s3, 1.0, bitmap, kind, b, sx, sy, sz, &ph);
                }
            }
        } else if (kind == 1) {
            for (int h = 0; h < g->n_holes_l0; ++h) {
                const float* hp = g->coll + 100 + 9 * h;
                double qsw[3] = {hp[0], ((double)hp[1] + (double)hp[2]) / 2.0 + g->F, hp[3]};
                OPS(20); /* hole centre, distance, the virtual-disc weight */
                scatter_body(g, &bd, d[i], p, qsw, 1.0, bitmap, 1, h, sx, sy, sz, &ph);
            }
        } else {
            for (int64_t k = 0; k < np; ++k)
                scatter_body(g, &bd, d[i], p, pts + 4 * k, 1.0, NULL, 2, -1, sx, sy, sz, &ph);
        }
    }
    g_count_ops = 0;
    free(pts);
    if (terms) *terms = g_terms;
    return g_terms > 0 ? g_ops / g_terms : 0.0;
}
