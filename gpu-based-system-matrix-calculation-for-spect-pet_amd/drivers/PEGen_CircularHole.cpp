// PEGen_CircularHole: photoelectric system matrix of all rotations.
//
// Same command line and files as the reference driver
// (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:31-155):
//   ./PEGen_CircularHole [-cuda GPU_ID]
// reads Params_{Collimator,Detector,Image,Physics}.dat from the working
// directory, sets Params_Image[20] to each rotation index in turn, calls
// PESysMatGen() and writes PE_SysMat_shift_<x>_<y>_<z>.sysmat, FP32
// [rotation][bar][voxel].
// Additions: -sub SX SY SZ (sub-cells per bar, default 8 16 8) and -fp64
// (all-FP64 kernels); they set SMG_PE_SUBSAMPLES / SMG_PRECISION.  The file is
// written one rotation at a time (the reference keeps all rotations in host
// memory), so its size is not bounded by host memory.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/smg_reference.h"
#include "driver_common.hpp"

static void usage(const char* argv0) {
    std::printf("Usage: %s [-cuda GPU_ID] [-sub SX SY SZ] [-fp64]\n", argv0);
}

int main(int argc, char* argv[]) {
    int cuda_id = 0;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-cuda") == 0 && i + 1 < argc) {
            cuda_id = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-sub") == 0 && i + 3 < argc) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%s,%s,%s", argv[i + 1], argv[i + 2], argv[i + 3]);
            setenv("SMG_PE_SUBSAMPLES", buf, 1);
            i += 3;
        } else if (std::strcmp(argv[i], "-fp64") == 0) {
            setenv("SMG_PRECISION", "fp64", 1);
        } else if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
            usage(argv[0]);
            return 0;
        } else {
            std::fprintf(stderr, "Unknown parameter or missing argument: %s\n", argv[i]);
            usage(argv[0]);
            return EXIT_FAILURE;
        }
    }
    drv::Params p = drv::read_params();
    drv::print_layers(p);
    drv::Dims d = drv::dims_of(p);
    std::printf("FOV dimension: %d %d %d, %d bars, %d rotations\n", d.nx, d.ny, d.nz, d.n_det,
                d.n_rot);
    // one rotation block in memory at a time; block r is written while r + 1 is computed
    drv::BlockWriter out(drv::sysmat_name("PE_SysMat", p), d.block());
    for (int r = 0; r < d.n_rot; ++r) {
        std::printf("rotation %d\n", r);
        p.img[20] = (float)r;
        int q = PESysMatGen(p.coll.data(), p.det.data(), p.img.data(), out.buffer(), cuda_id);
        std::printf("numImagebin = %d\n", q);
        out.commit();
    }
    out.close();
    return 0;
}
