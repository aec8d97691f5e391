// ScatterGen_Crystal: primary-Compton scatter between crystals (reference
// driver: ScatterGen_Crystal/ScatterGen_Crystal.cpp:29-216).
//   ./ScatterGen_Crystal -PE <PE matrix> -GeoCrystal <bitmap> -cuda <id>
// Always writes Scatter_Crystal_SysMat_shift_*; writes SysMat_shift_*
// (PE + scatter) if Params_Physics[3] == 1.
#include "../../include/smg_reference.h"
#include "scatter_driver.hpp"

int main(int argc, char* argv[]) {
    drv::help_exit(argc, argv, 1);
    drv::Params p = drv::read_params();
    drv::print_layers(p);
    drv::ScatterArgs a = drv::parse_scatter_args(argc, argv, p, 1);
    drv::Dims d = drv::dims_of(p);
    drv::scatter_rotations(p, d, a, "Scatter_Crystal_SysMat", true, "SysMat",
                           [&](int, float* pe_drive, float* pe_add, float* dst, float* sum) {
                               int q = scatter_and_sum(1, nullptr, p.det.data(), p.img.data(), p.phys.data(),
                                                       pe_drive, pe_add, a.geo_crystal.c_str(), nullptr, dst,
                                                       sum, a.cuda_id);
                               std::printf("numImagebin = %d\n", q);
                           });
    return 0;
}
