// ScatterGen_Collimator: primary-Compton scatter at the collimator holes, the
// r = 0.2 mm variant (reference driver: ScatterGen_Collimator/
// ScatterGen_Collimator.cpp:29-220).
//   ./ScatterGen_Collimator -PE <PE matrix> -GeoCollimator <bitmap> -cuda <id>
// Always writes Scatter_SysMat_shift_*; writes SysMat_shift_* (PE + scatter)
// if Params_Physics[3] == 1.
#include "../../include/smg_reference.h"
#include "scatter_driver.hpp"

int main(int argc, char* argv[]) {
    drv::help_exit(argc, argv, 2);
    drv::Params p = drv::read_params();
    drv::print_layers(p);
    drv::ScatterArgs a = drv::parse_scatter_args(argc, argv, p, 2);
    drv::Dims d = drv::dims_of(p);
    drv::scatter_rotations(p, d, a, "Scatter_SysMat", true, "SysMat",
                           [&](int, float* pe_drive, float* pe_add, float* dst, float* sum) {
                               int q = scatter_and_sum(2, p.coll.data(), p.det.data(), p.img.data(),
                                                       p.phys.data(), pe_drive, pe_add, nullptr,
                                                       a.geo_coll.c_str(), dst, sum, a.cuda_id);
                               std::printf("numImagebin = %d\n", q);
                           });
    return 0;
}
