// Shared pieces of the four drivers: parameter files, sizes, matrix files.
//
// Same file contract as the reference drivers
// (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:33-56): the four
// Params_*.dat files in the working directory, 80000/80000/100/100 raw
// float32 values.  A missing or short file is an error (exit status 1); the
// reference read them unchecked.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace drv {

struct Params {
    std::vector<float> coll = std::vector<float>(80000, 0.0f);
    std::vector<float> det = std::vector<float>(80000, 0.0f);
    std::vector<float> img = std::vector<float>(100, 0.0f);
    std::vector<float> phys = std::vector<float>(100, 0.0f);
};

inline void read_floats(const char* name, std::vector<float>& dst) {
    FILE* f = std::fopen(name, "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", name);
        std::exit(EXIT_FAILURE);
    }
    size_t got = std::fread(dst.data(), sizeof(float), dst.size(), f);
    std::fclose(f);
    if (got != dst.size()) {
        std::fprintf(stderr, "%s holds %zu floats, %zu expected\n", name, got, dst.size());
        std::exit(EXIT_FAILURE);
    }
}

inline Params read_params() {
    Params p;
    read_floats("Params_Collimator.dat", p.coll);
    read_floats("Params_Detector.dat", p.det);
    read_floats("Params_Image.dat", p.img);
    read_floats("Params_Physics.dat", p.phys);
    return p;
}

struct Dims {
    int n_det, nx, ny, nz, n_rot;
    long long n_vox() const { return (long long)nx * ny * nz; }
    size_t block() const { return (size_t)n_det * (size_t)n_vox(); }
};

inline Dims dims_of(const Params& p) {
    Dims d;
    d.n_det = (int)std::floor(p.det[0] + 0.001f);
    d.nx = (int)std::floor(p.img[0] + 0.001f);
    d.ny = (int)std::floor(p.img[1] + 0.001f);
    d.nz = (int)std::floor(p.img[2] + 0.001f);
    d.n_rot = (int)std::floor(p.img[6] + 0.001);
    if (d.n_rot < 1) d.n_rot = 1;
    return d;
}

inline void print_layers(const Params& p) {
    int layers = (int)std::floor(p.coll[0] + 0.001f);
    for (int m = 0; m < layers; ++m) {
        const float* L = p.coll.data() + (m + 1) * 10;
        std::printf("collimator layer %d: %g holes, %g x %g x %g mm, at %g mm, mu %g/%g/%g per mm\n",
                    m, L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7]);
    }
    std::printf("FOV centre to first collimator = %g mm\n", p.img[11]);
}

// "<prefix>_shift_%f_%f_%f.sysmat" (PEGen_CircularHole.cpp:141-142)
inline std::string sysmat_name(const char* prefix, const Params& p) {
    char buf[2048];
    std::snprintf(buf, sizeof buf, "%s_shift_%f_%f_%f.sysmat", prefix, p.img[8], p.img[9], p.img[10]);
    return buf;
}

inline void write_matrix(const std::string& name, const std::vector<float>& m) {
    FILE* f = std::fopen(name.c_str(), "wb+");
    if (!f || std::fwrite(m.data(), sizeof(float), m.size(), f) != m.size()) {
        std::fprintf(stderr, "cannot write %s\n", name.c_str());
        std::exit(EXIT_FAILURE);
    }
    std::fclose(f);
    std::printf("%s written\n", name.c_str());
}

inline void read_matrix(const std::string& name, std::vector<float>& m) {
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open PE matrix %s\n", name.c_str());
        std::exit(EXIT_FAILURE);
    }
    size_t got = std::fread(m.data(), sizeof(float), m.size(), f);
    std::fclose(f);
    if (got != m.size()) {
        std::fprintf(stderr, "%s holds %zu floats, %zu expected\n", name.c_str(), got, m.size());
        std::exit(EXIT_FAILURE);
    }
}

}  // namespace drv
