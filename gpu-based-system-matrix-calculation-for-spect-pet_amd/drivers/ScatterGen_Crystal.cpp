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
    std::vector<float> pe(d.block() * (size_t)d.n_rot), out(d.block() * (size_t)d.n_rot, 0.0f);
    drv::read_matrix(a.pe_file, pe);
    for (int r = 0; r < d.n_rot; ++r) {
        std::printf("rotation %d\n", r);
        p.img[20] = (float)r;
        float* pe_block = pe.data() + (a.pe_per_rotation ? (size_t)r * d.block() : 0);
        int q = scatter(p.det.data(), p.img.data(), p.phys.data(), pe_block, a.geo_crystal.c_str(),
                        out.data() + (size_t)r * d.block(), a.cuda_id);
        std::printf("numImagebin = %d\n", q);
    }
    drv::write_matrix(drv::sysmat_name("Scatter_Crystal_SysMat", p), out);
    if (p.phys[3] == 1) {
        std::vector<float> sum(out.size());
        for (size_t i = 0; i < out.size(); ++i) sum[i] = pe[i] + out[i];
        drv::write_matrix(drv::sysmat_name("SysMat", p), sum);
    }
    return 0;
}
