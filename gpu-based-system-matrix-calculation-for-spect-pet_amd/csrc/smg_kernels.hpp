// Launch descriptors of the gfx950 kernels (host and device side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "smg_pe_body.hpp"
#include "smg_scatter_body.hpp"

namespace smg {

// LDS capacity of one wave's candidate lists (bars, holes): with the rest of
// a PE wave's LDS 10 KB, so 16 one-wave blocks fit a CU.  A wave whose
// candidates exceed these is not truncated: it queues its work item, and a
// second kernel redoes the queued items with lists in global memory (PeSpill).
constexpr int kCapBars = 448;
constexpr int kCapHoles = 256;

// Waves of a spill kernel (one list slot each).
constexpr int kSpillWaves = 1024;

// Work items whose candidates overflowed the LDS lists.  Each PE-family
// launch clears counters[2], its kernel appends the work item of every
// overflowing wave to queue (room for one entry per block of the launch),
// and the launch's spill kernel (kSpillWaves waves or fewer) redoes the
// queued items, wave b with the lists in slot b of slots; cap_bars and
// cap_holes are the geometry's bar and hole counts, so those lists cannot
// overflow.  The spilled items are added to counters[1].
struct PeSpill {
    int* queue = nullptr;
    int64_t queue_cap = 0;
    unsigned char* slots = nullptr;
    int64_t slot_bytes = 0;
    int n_slots = 0;
    int cap_bars = 0, cap_holes = 0;
};

// Bytes of one spill slot: masks, adjoint sums and indices of every bar and
// hole, and the face records of the first narrowed bars, in 256-byte units.
inline int64_t pe_spill_slot_bytes(int n_det, int n_holes) {
    const int64_t b = n_det > 0 ? n_det : 1, h = n_holes > 0 ? n_holes : 1;
    const int64_t bytes = b * (8 + 8 + 2 + 2) + h * (8 + 2 + 2) + 32 * (7 * 4 + 1) + (b + 63) / 64 * 8;
    return (bytes + 255) / 256 * 256;
}

struct PeLaunch {
    PeGeom geom;
    float* dst;                   // [N][V] block of this rotation
    unsigned long long* counters; // [0] FP64 fallback rays, [1] work items redone with
                                  // spill lists, [2] spill queue length of the launch
    double cphi, sphi;            // rotation of the voxel grid about z
    int sub[3];                   // sub-cells per bar (x, y, z)
    int tile[3];                  // voxel tile of one wave, product 64
    int tiles[3];                 // number of tiles per axis
    int bar0;                     // first bar of the launch (point weights)
    float* pts;                   // point weights [bar - bar0][cell][V]
    int cell_slices;              // point weights: waves per (bar, tile), each a range of cells
    const CellCtx* cells;         // sub-cell table [bar - cells_bar0][cell] (cell_table), or
    int cells_bar0;               // nullptr: the kernel forms each sub-cell itself
    // point weights: the detector rows that will use them, 5 doubles each (bar
    // centre, bounding radius, energy-window cosine cut); a point whose terms
    // are below the window for every row and voxel of the tile gets weight 0
    // without its ray.  n_use_rows = 0: every weight is formed.
    const double* use_rows = nullptr;
    int n_use_rows = 0;
    PeSpill spill;                // lists of the waves that overflow LDS
    // a hole is elliptical: the FP32 matrix and projection launches run
    // their kernels with the elliptic-cylinder chord compiled in (the FP64
    // kernels always carry it; the other FP32 PE launches do not)
    bool ell = false;
};

// Matrix-free projections with the PE matrix of A's rotation: y = A x
// (x: V floats, y: N floats) and x = A^T y.  partial: project_scratch(A, back)
// doubles.
size_t project_scratch(const PeLaunch& a, bool back);
hipError_t launch_pe_project(const PeLaunch& a, bool exact, const float* x, double* partial, float* y,
                             hipStream_t stream);
hipError_t launch_pe_backproject(const PeLaunch& a, bool exact, const float* y, double* partial, float* x,
                                 hipStream_t stream);

// Sub-cell table of bars bar0 .. bar0 + n_bars - 1 for sub-cells sub[3], in
// the PE kernels' cell order: the same values those kernels would compute.
hipError_t launch_cell_table(const PeGeom& geom, int bar0, int n_bars, const int sub[3], CellCtx* out,
                             hipStream_t stream);

// Tangent (forward-mode derivative) of the PE matrix along a direction v over
// the adjoint's parameters ([n_layers layer mu][N bar mu_total][N bar mu_pe]):
// dA[d][v] = sum_k v_k dA[d][v]/dmu_k, and A itself when a_out is set.
struct PeTangent {
    const double* v;   // direction (device)
    float* d_out;      // [N][V]
    float* a_out;      // [N][V] or nullptr
};
hipError_t launch_pe_tangent(const PeLaunch& a, const PeTangent& t, bool exact, hipStream_t stream);

hipError_t launch_pe_forward(const PeLaunch& a, bool exact, hipStream_t stream);
// The rows (bars) bar0 .. bar0 + n_bars - 1 of the same matrix.
hipError_t launch_pe_forward_rows(const PeLaunch& a, bool exact, int bar0, int n_bars, hipStream_t stream);
// Crystal grid scatterers: first-leg weights of the sub-cell centres of bars
// bar0 .. bar0 + n_bars - 1 (sub-cells = grid cells) into a.pts.
hipError_t launch_pe_points(const PeLaunch& a, bool exact, int bar0, int n_bars, hipStream_t stream);

// Adjoint scratch: one row of `stride` doubles per (bar d, tile group):
// [n_det mu_total partials][n_layers layer partials][mu_pe of d][loss].
struct PeGradScratch {
    double* rows;
    int64_t stride;
    int groups;       // waves per bar
    int cell_slices;  // sub-cell ranges per tile (work items per bar: tiles x cell_slices)
    const float* w;   // [N][V] weights of L = sum w A
};

hipError_t launch_pe_backward(const PeLaunch& a, const PeGradScratch& s, bool exact, double* grad,
                              hipStream_t stream);

struct BitmapLaunch {
    const Sphere<float>* spheres_f;   // bars
    const Sphere<double>* spheres_d;
    const Sphere<float>* holes_f;     // layer-0 hole centres
    const Sphere<double>* holes_d;
    uint32_t* words;
    int64_t n_words;
    int n_det, n_holes;
};

hipError_t launch_bitmap(const BitmapLaunch& a, bool collimator, hipStream_t stream);

struct ScatterLaunch {
    PeGeom geom;
    const ScatterPt<double>* pts_d;  // scatter points
    int n_pts;
    int kind;                        // 0: crystal scatter, 1: collimator scatter,
                                     // 2: collimator scatter from grid points,
                                     // 3: crystal scatter from grid points
    int variant;                     // collimator: 0 combined (r = 0.1), 1 standalone (r = 0.2)
    const float* pe;                 // crystal: PE block [N][V]
    const uint32_t* bitmap;          // crystal or collimator bitmap
    PairTerm* pair;                  // [N][n_pts][cells]
    float* dst;                      // [N][V]
    unsigned long long* counters;    // [0] FP64 fallback terms
    double cphi, sphi;
    int sub[3];                      // detection sub-cells
    double e0, win_lo, win_hi;       // photopeak energy, energy window
    double kn_norm;                  // KN sum over [0, pi]
    double coll_ratio;               // mu_compton / mu_total of collimator layer 0
    int same_window, compton_on;
    const MuTables* tables;          // mu(E) tables, or nullptr for the power laws
    const MuTablesF* tables_f;       // their FP32 form (when tables is set)
    const float* w1;                 // grid: first-leg weights [n_pts][V] of this chunk
    int accumulate;                  // add to dst instead of writing it
    const int* rows;                 // detector rows to compute (dst row r = bar rows[r]),
    int n_rows;                      // or nullptr for all bars
    int pts_per_bar, pt_bar0;        // kind 3: point s lies in bar pt_bar0 + s / pts_per_bar
    double* partial;                 // [slices][rows][V] per-slice sums when the scatter points
                                     // are split over several waves per element
    const double* kn_table;          // KN prefix table of e0 (FP32 path), or nullptr
    const double* row_cut;           // [rows] energy-window cut of each output row
                                     // (window_reject_cos, no margin; launch_row_cuts)
};

// KN prefix table of e0 ((kKnPhase + 1) x kKnSteps doubles) into tab.
hipError_t launch_kn_table(double e0, double* tab, hipStream_t stream);

// Grid scatterers: first-leg weights w1[s][v] = w_s / (4 pi L^2) exp(-att) of
// the chunk's scatter points (w_s = dv mu_c of the point's layer, in its dy).
hipError_t launch_first_leg(const ScatterLaunch& a, float* w1, hipStream_t stream);

// The energy-window cut of each output row into cut[rows], read by the pair
// terms and scatter kernels of the same launch (a.row_cut).
hipError_t launch_row_cuts(const ScatterLaunch& a, double* cut, hipStream_t stream);
hipError_t launch_pair_terms(const ScatterLaunch& a, hipStream_t stream);
// out[i] = a[i] + b[i], i < n (the PE + scatter block).
hipError_t launch_add_blocks(const float* a, const float* b, float* out, size_t n, hipStream_t stream);
hipError_t launch_scatter(const ScatterLaunch& a, bool exact, hipStream_t stream);
// Doubles launch_scatter needs in `partial` for this launch (0: one wave per
// element covers every scatter point, no partial sums).
size_t scatter_partial_size(const ScatterLaunch& a);

}  // namespace smg
