// ScatterGen_CircularHole: primary-Compton scatter in the crystals and at the
// collimator holes, summed (reference driver:
// ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:38-232).
//   ./ScatterGen_CircularHole -PE <PE matrix> -GeoCrystal <bitmap>
//                             -GeoCollimator <bitmap> -cuda <id>
// Writes Scatter_SysMat_shift_* if Params_Physics[2] == 1 and
// SysMat_withScatter_shift_* (PE + scatter) if Params_Physics[3] == 1.
#include "../../include/smg_reference.h"
#include "scatter_driver.hpp"

int main(int argc, char* argv[]) {
    drv::help_exit(argc, argv, 3);
    drv::Params p = drv::read_params();
    drv::print_layers(p);
    drv::ScatterArgs a = drv::parse_scatter_args(argc, argv, p, 3);
    drv::Dims d = drv::dims_of(p);
    drv::scatter_rotations(p, d, a, "Scatter_SysMat", p.phys[2] == 1, "SysMat_withScatter",
                           [&](int, float* pe_drive, float* pe_add, float* dst, float* sum) {
                               int q = scatter_and_sum(3, p.coll.data(), p.det.data(), p.img.data(),
                                                       p.phys.data(), pe_drive, pe_add, a.geo_crystal.c_str(),
                                                       a.geo_coll.c_str(), dst, sum, a.cuda_id);
                               std::printf("numImagebin = %d\n", q);
                           });
    return 0;
}
