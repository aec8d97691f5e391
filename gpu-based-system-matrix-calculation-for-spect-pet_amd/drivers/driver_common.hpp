// Shared pieces of the four drivers: parameter files, sizes, matrix files.
//
// Same file contract as the reference drivers
// (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:33-56): the four
// Params_*.dat files in the working directory, 80000/80000/100/100 raw
// float32 values.  A missing or short file is an error (exit status 1); the
// reference read them unchecked.
#pragma once

#include <sys/types.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
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

// Sequential writer of a [rotation][bar][voxel] matrix file, one rotation
// block at a time, so a driver holds two blocks in memory whatever the number
// of rotations (the reference drivers keep the whole matrix, R x N x V floats,
// and write it at the end).  A committed block is written by a background
// thread while the caller fills the other buffer; the file is built as
// "<name>.part" and renamed to <name> once complete.  The bytes are those of
// the whole-matrix write.
class BlockWriter {
  public:
    BlockWriter(const std::string& name, size_t block)
        : name_(name), part_(name + ".part"), block_(block) {
        f_ = std::fopen(part_.c_str(), "wb");
        if (!f_) fail("cannot write");
        buf_[0].assign(block, 0.0f);
        buf_[1].assign(block, 0.0f);
    }
    BlockWriter(const BlockWriter&) = delete;
    BlockWriter& operator=(const BlockWriter&) = delete;
    ~BlockWriter() {
        if (th_.joinable()) th_.join();
        if (f_) std::fclose(f_);
    }
    // the buffer of the next block
    float* buffer() { return buf_[cur_].data(); }
    // hand the buffer's block to the writer thread (after the previous write)
    void commit() {
        if (th_.joinable()) th_.join();
        if (!ok_) fail("cannot write");
        const float* b = buf_[cur_].data();
        th_ = std::thread([this, b] { ok_ = std::fwrite(b, sizeof(float), block_, f_) == block_; });
        cur_ ^= 1;
    }
    void close() {
        if (th_.joinable()) th_.join();
        if (!ok_) fail("cannot write");
        if (std::fclose(f_) != 0) {
            f_ = nullptr;
            fail("cannot write");
        }
        f_ = nullptr;
        if (std::rename(part_.c_str(), name_.c_str()) != 0) fail("cannot rename to");
        std::printf("%s written\n", name_.c_str());
    }

  private:
    [[noreturn]] void fail(const char* what) {
        std::fprintf(stderr, "%s %s\n", what, name_.c_str());
        std::exit(EXIT_FAILURE);
    }
    std::string name_, part_;
    size_t block_;
    FILE* f_ = nullptr;
    std::vector<float> buf_[2];
    int cur_ = 0;
    bool ok_ = true;  // set by the writer thread, read after joining it
    std::thread th_;
};

// Rotation blocks of a [rotation][bar][voxel] matrix file, read one at a time;
// the file must hold at least n_blocks blocks (the reference reads them all).
class BlockReader {
  public:
    BlockReader(const std::string& name, size_t block, int n_blocks) : name_(name), block_(block) {
        f_ = std::fopen(name.c_str(), "rb");
        if (!f_) {
            std::fprintf(stderr, "cannot open PE matrix %s\n", name.c_str());
            std::exit(EXIT_FAILURE);
        }
        const size_t want = block * (size_t)n_blocks;
        size_t have = 0;
        if (fseeko(f_, 0, SEEK_END) == 0) have = (size_t)ftello(f_) / sizeof(float);
        if (have < want) {
            std::fprintf(stderr, "%s holds %zu floats, %zu expected\n", name.c_str(), have, want);
            std::exit(EXIT_FAILURE);
        }
    }
    BlockReader(const BlockReader&) = delete;
    BlockReader& operator=(const BlockReader&) = delete;
    ~BlockReader() { std::fclose(f_); }
    void read(int r, float* dst) {
        if (fseeko(f_, (off_t)r * (off_t)block_ * (off_t)sizeof(float), SEEK_SET) != 0 ||
            std::fread(dst, sizeof(float), block_, f_) != block_) {
            std::fprintf(stderr, "cannot read rotation %d of %s\n", r, name_.c_str());
            std::exit(EXIT_FAILURE);
        }
    }

  private:
    std::string name_;
    size_t block_;
    FILE* f_ = nullptr;
};

}  // namespace drv
