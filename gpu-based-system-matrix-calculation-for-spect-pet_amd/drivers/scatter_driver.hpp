// Command line and rotation loop shared by the three scatter drivers
// (ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:38-232,
// ScatterGen_Collimator/ScatterGen_Collimator.cpp:29-220,
// ScatterGen_Crystal/ScatterGen_Crystal.cpp:29-216).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "driver_common.hpp"

namespace drv {

struct ScatterArgs {
    std::string pe_file, geo_crystal, geo_coll;
    int cuda_id = 0;
    bool pe_per_rotation = false;
};

inline void print_scatter_usage(const char* argv0, int which) {
    std::printf("Usage: %s [-PE PE_SysMat_path]%s%s [-cuda GPU_ID] [-PErot] [-sub SX SY SZ] [-fp64]\n",
                argv0, (which & 1) ? " [-GeoCrystal GeometryRelationShip_Crystal2Crystal_path]" : "",
                (which & 2) ? " [-GeoCollimator GeometryRelationShip_Collimator2Crystal_path]" : "");
}

// -h / --help answers before the parameter files are read.
inline void help_exit(int argc, char* argv[], int which) {
    for (int i = 1; i < argc; ++i)
        if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
            print_scatter_usage(argv[0], which);
            std::exit(0);
        }
}

// accepts: which = 1 crystal, 2 collimator, 3 both
inline ScatterArgs parse_scatter_args(int argc, char* argv[], Params& p, int which) {
    ScatterArgs a;
    a.geo_crystal = "GeometryRelationShip_Crystal2Crystal";
    a.geo_coll = "GeometryRelationShip_Collimator2Crystal";
    auto usage = [&]() {
        std::printf("Usage: %s [-PE PE_SysMat_path]%s%s [-cuda GPU_ID] [-PErot] [-sub SX SY SZ] [-fp64]\n",
                    argv[0],
                    (which & 1) ? " [-GeoCrystal GeometryRelationShip_Crystal2Crystal_path]" : "",
                    (which & 2) ? " [-GeoCollimator GeometryRelationShip_Collimator2Crystal_path]" : "");
    };
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-PE") == 0 && i + 1 < argc) {
            a.pe_file = argv[++i];
        } else if ((which & 1) && std::strcmp(argv[i], "-GeoCrystal") == 0 && i + 1 < argc) {
            a.geo_crystal = argv[++i];
            p.phys[8] = 0;  // load instead of compute
        } else if ((which & 2) && std::strcmp(argv[i], "-GeoCollimator") == 0 && i + 1 < argc) {
            a.geo_coll = argv[++i];
            p.phys[9] = 0;
        } else if (std::strcmp(argv[i], "-cuda") == 0 && i + 1 < argc) {
            a.cuda_id = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-PErot") == 0) {
            // use each rotation's own PE block (the reference passes the
            // rotation-0 block for every rotation)
            a.pe_per_rotation = true;
        } else if (std::strcmp(argv[i], "-sub") == 0 && i + 3 < argc) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%s,%s,%s", argv[i + 1], argv[i + 2], argv[i + 3]);
            setenv("SMG_SCATTER_SUBSAMPLES", buf, 1);
            i += 3;
        } else if (std::strcmp(argv[i], "-fp64") == 0) {
            setenv("SMG_PRECISION", "fp64", 1);
        } else if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
            usage();
            std::exit(0);
        } else {
            std::fprintf(stderr, "Unknown parameter or missing argument: %s\n", argv[i]);
            usage();
            std::exit(EXIT_FAILURE);
        }
    }
    if (a.pe_file.empty()) a.pe_file = sysmat_name("PE_SysMat", p);
    return a;
}

}  // namespace drv
