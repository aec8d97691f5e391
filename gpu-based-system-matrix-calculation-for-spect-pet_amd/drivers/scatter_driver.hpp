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

// Rotation loop of the scatter drivers with one rotation block of each matrix
// in memory.  compute(r, pe_drive, pe_add, scatter_dst, sum_dst) fills rotation
// r's scatter block from the PE block that drives it (rotation 0's for every
// rotation, as the reference drivers pass it, or rotation r's with -PErot) and,
// when sum_dst is not null, the PE (rotation r) + scatter block, the sum the
// reference drivers form at the end; both blocks come from the GPU.  Writes
// scatter_name when write_scatter, and sum_name when Params_Physics[3] == 1.
template <class Compute>
void scatter_rotations(Params& p, const Dims& d, const ScatterArgs& a, const char* scatter_name,
                       bool write_scatter, const char* sum_name, Compute compute) {
    std::printf("Photon Electric SysMat: %s\n", a.pe_file.c_str());
    BlockReader pe_file(a.pe_file, d.block(), d.n_rot);
    const bool write_sum = p.phys[3] == 1;
    std::vector<float> pe0(d.block()), pe_r(d.block()), out;
    pe_file.read(0, pe0.data());
    BlockWriter* ws = write_scatter ? new BlockWriter(sysmat_name(scatter_name, p), d.block()) : nullptr;
    BlockWriter* wsum = write_sum ? new BlockWriter(sysmat_name(sum_name, p), d.block()) : nullptr;
    if (!ws && !wsum) out.assign(d.block(), 0.0f);  // computed and not written, as the reference does
    for (int r = 0; r < d.n_rot; ++r) {
        std::printf("rotation %d\n", r);
        p.img[20] = (float)r;
        float* pe_here = pe0.data();  // rotation r's PE block (non-const: the reference signatures)
        if (r > 0 && (a.pe_per_rotation || write_sum)) {
            pe_file.read(r, pe_r.data());
            pe_here = pe_r.data();
        }
        float* dst = ws ? ws->buffer() : (wsum ? nullptr : out.data());
        compute(r, a.pe_per_rotation ? pe_here : pe0.data(), pe_here, dst, wsum ? wsum->buffer() : nullptr);
        if (wsum) wsum->commit();
        if (ws) ws->commit();
    }
    if (ws) ws->close();
    if (wsum) wsum->close();
    delete ws;
    delete wsum;
}

}  // namespace drv
