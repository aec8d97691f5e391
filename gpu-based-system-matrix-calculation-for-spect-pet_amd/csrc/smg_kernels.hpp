// Launch descriptors of the gfx950 kernels (host and device side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "smg_pe_body.hpp"

namespace smg {

// LDS capacity of one wave's candidate lists (bars, holes).  A wave whose
// conservative candidate count exceeds these is reported, not truncated.
constexpr int kCapBars = 512;
constexpr int kCapHoles = 256;

struct PeLaunch {
    PeGeom geom;
    float* dst;                   // [N][V] block of this rotation
    unsigned long long* counters; // [0] FP64 fallback rays, [1] overflowed waves
    double cphi, sphi;            // rotation of the voxel grid about z
    int sub[3];                   // sub-cells per bar (x, y, z)
    int tile[3];                  // voxel tile of one wave, product 64
    int tiles[3];                 // number of tiles per axis
};

hipError_t launch_pe_forward(const PeLaunch& a, bool exact, hipStream_t stream);

}  // namespace smg
