// Device-resident geometry of the system-matrix generator.
//
// The reference reads every geometric quantity straight out of the raw float
// parameter arrays inside each thread (PEGen_RayTracing_CircularHole/
// PESysMatGen.cu:958-1075).  Here the host parses the arrays once into small
// fixed-stride records (one 64-byte record per bar, one per hole), in a
// float copy and an exact double copy, so a wave can fetch one record with a
// single scalar load.
#pragma once

#include <cstdint>

namespace smg {

// Capacities of the parameter arrays (80000 floats per array).
constexpr int kMaxLayers = 9;
constexpr int kMaxHoles = (80000 - 100) / 9;
constexpr int kMaxBars = (80000 - 1) / 12;

template <class T>
struct Bar {
    T cx, cy, cz;   // centre; cy includes the FOV-to-collimator distance F
    T w, t, h;      // width (x), thickness (y), height (z)
    T mut, mupe, muc, res;
    T cb, sb;       // cos(beta), sin(beta); beta = rotation about y
    T rdiag;        // sqrt(w^2 + h^2 + t^2): culling uses rdiag / (2 D)
    T beta;
    T pe_frac;      // mu_pe / mu_total
    T pad1;
};

template <class T>
struct Hole {
    T xc, y1, y2, zc;  // y1, y2 include F
    T r, mu;           // radius, mu_total of the hole filling
    T ymid;            // (y1 + y2) / 2
    T crit;            // sqrt(4 r^2 + (y1 - y2)^2) / 2: culling uses crit / D
    T mupe, muc;       // photoelectric and Compton mu of the filling
    T erx, erz;        // elliptical hole: semi-axes along x' and z' (erx = 0: circle of radius r;
                       // otherwise r = max(erx, erz), the bounding radius)
    T ec, es;          // cos and sin of the turn theta of (x', z') from (x, z)
};

template <class T>
struct Layer {
    T lo[3], hi[3];   // plate box (world frame, F included)
    T thick, mu;      // thickness (y) and mu_total of the plate
    T mupe, muc;      // photoelectric and Compton mu of the plate
};

// Bounding sphere of a bar, or centre of a layer-0 hole, in the raw
// parameter frame (no F added): the bitmaps use these positions
// (ScatterGen_RayTracing_CircularHole/scatter.cu:1465-1467, :1963-1965).
template <class T>
struct Sphere {
    T x, y, z, r;
};

// A scatter point: bar centre (crystal scatter) or layer-0 hole centre
// (collimator scatter), world frame with F; dy = y2 - y1 of the hole.
template <class T>
struct ScatterPt {
    T x, y, z, dy;
};

// Axis-aligned world box (centre, half sizes) around kGroup consecutive bars
// or holes: the candidate scan of the PE kernel tests a group's box before its
// members.
constexpr int kGroup = 64;
struct GroupBox {
    float c[3], h[3];
};

struct LayerRange {
    int32_t first, count;  // holes of this layer in the global hole table
};

// Everything a kernel needs to decode voxels and address the parameter
// records.  Passed by value as a kernel argument.
struct GeomDesc {
    int32_t n_det, n_layers, n_holes;
    int32_t nx, ny, nz;
    int64_t n_vox;
    double wx, wy, wz;        // voxel widths
    double sx, sy, sz;        // FOV shift
    double dphi;              // angle per rotation
    double fov2coll;          // F
    LayerRange ranges[kMaxLayers];
};

}  // namespace smg
