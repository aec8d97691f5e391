// Host-side parsing of the four parameter arrays into device records.
//
// Integer fields are stored as floats and read with floor(x + small eps), as
// the reference does (PEGen_RayTracing_CircularHole/PESysMatGen.cu:961-964,
// :1377-1383); positions get the FOV-to-collimator distance F added to y
// (PESysMatGen.cu:1059, :1160-1161).  Shared by the C ABI and the host unit
// tests of the lane arithmetic.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "smg_geometry.hpp"

namespace smg {

// Host copy of the parsed geometry.
struct HostGeom {
    GeomDesc g{};
    std::vector<Bar<float>> bars_f;
    std::vector<Bar<double>> bars_d;
    std::vector<Hole<float>> holes_f;
    std::vector<Hole<double>> holes_d;
    std::vector<Layer<float>> layers_f;
    std::vector<Layer<double>> layers_d;
    std::vector<GroupBox> bar_groups, hole_groups;
    int n_rot = 1;
    int n_holes_l0 = 0;
    std::vector<float> coll, det, img, phys;  // raw arrays (first 80000/80000/100/100)
};

// Error codes match smg_status.
constexpr int kBadParam = 1;
constexpr int kCapacity = 2;

inline int perr(std::string& why, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    why = buf;
    return code;
}

// World boxes around each run of kGroup bars (turned by beta about y) and
// holes (cylinders along y), padded by 0.01 for the FP32 rounding of centres.
inline void make_group_boxes(HostGeom& hg) {
    auto build = [](int n, auto box_of, std::vector<GroupBox>& out) {
        out.assign((n + kGroup - 1) / kGroup, GroupBox{});
        for (int gi = 0; gi < (int)out.size(); ++gi) {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int i = gi * kGroup; i < n && i < (gi + 1) * kGroup; ++i) {
                double c[3], h[3];
                box_of(i, c, h);
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], c[k] - h[k]);
                    hi[k] = std::max(hi[k], c[k] + h[k]);
                }
            }
            for (int k = 0; k < 3; ++k) {
                out[gi].c[k] = (float)(0.5 * (lo[k] + hi[k]));
                out[gi].h[k] = (float)(0.5 * (hi[k] - lo[k]) + 0.01 + 1e-5 * std::fabs(0.5 * (lo[k] + hi[k])));
            }
        }
    };
    build(hg.g.n_det, [&](int i, double c[3], double h[3]) {
        const Bar<double>& b = hg.bars_d[i];
        const double ac = std::fabs(b.cb), as = std::fabs(b.sb);
        c[0] = b.cx; c[1] = b.cy; c[2] = b.cz;
        h[0] = ac * 0.5 * b.w + as * 0.5 * b.h;
        h[1] = 0.5 * b.t;
        h[2] = as * 0.5 * b.w + ac * 0.5 * b.h;
    }, hg.bar_groups);
    build(hg.g.n_holes, [&](int i, double c[3], double h[3]) {
        const Hole<double>& o = hg.holes_d[i];
        c[0] = o.xc; c[1] = o.ymid; c[2] = o.zc;
        h[0] = o.r;
        h[1] = 0.5 * std::fabs(o.y2 - o.y1);
        h[2] = o.r;
    }, hg.hole_groups);
}

// True when the point (x, y, z) lies inside hole o (its ellipse or circle,
// strictly, and within its y span): the collimator grid leaves such points out.
inline bool in_hole(const Hole<double>& o, double x, double y, double z) {
    const double dx = x - o.xc, dz = z - o.zc;
    bool in;
    if (o.erx > 0.0) {
        const double u = dx * o.ec + dz * o.es, w = -dx * o.es + dz * o.ec;
        in = u * u / (o.erx * o.erx) + w * w / (o.erz * o.erz) < 1.0;
    } else {
        in = dx * dx + dz * dz < o.r * o.r;
    }
    return in && y >= o.y1 && y <= o.y2;
}

// Scatter points on a grid over the plates of the layers in layer_mask (bit
// m: layer m; the collimator grid-scatterer mode, an extension with no
// reference counterpart): per layer, n_k = max(1, round(extent / pitch))
// cells per axis, points at the cell centres, z outermost and x innermost,
// without the points inside a hole of that layer; layers in increasing
// order.  Each point's dy holds its weight factor, the cell volume times the
// layer's Compton mu.
inline void make_coll_grid(const HostGeom& hg, double pitch, unsigned layer_mask,
                           std::vector<ScatterPt<double>>& pts) {
    pts.clear();
    if (!(pitch > 0.0)) return;
    for (int m = 0; m < hg.g.n_layers; ++m) {
        if (!((layer_mask >> m) & 1u)) continue;
        const Layer<double>& L = hg.layers_d[m];
        int n[3];
        double pk[3];
        // extents as the parameters give them: width, thickness, height
        const double exts[3] = {L.hi[0] - L.lo[0], L.thick, L.hi[2] - L.lo[2]};
        for (int k = 0; k < 3; ++k) {
            n[k] = (int)std::floor(exts[k] / pitch + 0.5);
            if (n[k] < 1) n[k] = 1;
            pk[k] = exts[k] / n[k];
        }
        const double w = pk[0] * pk[1] * pk[2] * L.muc;
        const int h0 = hg.g.ranges[m].first, h1 = h0 + hg.g.ranges[m].count;
        for (int kz = 0; kz < n[2]; ++kz)
            for (int jy = 0; jy < n[1]; ++jy)
                for (int ix = 0; ix < n[0]; ++ix) {
                    const double x = L.lo[0] + (ix + 0.5) * pk[0], y = L.lo[1] + (jy + 0.5) * pk[1],
                                 z = L.lo[2] + (kz + 0.5) * pk[2];
                    bool inside = false;
                    for (int h = h0; h < h1 && !inside; ++h) inside = in_hole(hg.holes_d[h], x, y, z);
                    if (!inside) pts.push_back(ScatterPt<double>{x, y, z, w});
                }
    }
}

// The culling radius and the FP32 copy of hole h's record.
inline void hole_derived(HostGeom& hg, int h) {
    Hole<double>& o = hg.holes_d[h];
    o.crit = std::sqrt(o.r * o.r * 4 + (o.y1 - o.y2) * (o.y1 - o.y2)) / 2.0;
    Hole<float>& of = hg.holes_f[h];
    of.xc = (float)o.xc;
    of.y1 = (float)o.y1;
    of.y2 = (float)o.y2;
    of.zc = (float)o.zc;
    of.r = (float)o.r;
    of.mu = (float)o.mu;
    of.mupe = (float)o.mupe;
    of.muc = (float)o.muc;
    of.ymid = (float)o.ymid;
    of.crit = (float)o.crit;
    of.erx = (float)o.erx;
    of.erz = (float)o.erz;
    of.ec = (float)o.ec;
    of.es = (float)o.es;
}

// Parse the four parameter arrays the way the reference reads them.
inline int parse_geometry(const float* coll, const float* det, const float* img, const float* phys,
                          HostGeom& hg, std::string& why) {
    GeomDesc& g = hg.g;
    std::memset(&g, 0, sizeof g);
    hg.coll.assign(coll, coll + 80000);
    hg.det.assign(det, det + 80000);
    hg.img.assign(img, img + 100);
    hg.phys.assign(phys, phys + 100);
    // integer fields stored as floats (PESysMatGen.cu:961-964, :1377-1383)
    g.n_layers = (int)std::floor((double)coll[0] + 1e-6);
    g.n_det = (int)std::floor((double)det[0] + 1e-4);
    g.nx = (int)std::floor((double)img[0] + 1e-3);
    g.ny = (int)std::floor((double)img[1] + 1e-3);
    g.nz = (int)std::floor((double)img[2] + 1e-3);
    hg.n_rot = (int)std::floor((double)img[6] + 1e-3);
    if (hg.n_rot < 1) hg.n_rot = 1;
    if (g.n_layers < 0 || g.n_layers > kMaxLayers)
        return perr(why, kCapacity, "collimator layers %d outside [0, %d]", g.n_layers, kMaxLayers);
    if (g.n_det < 1 || g.n_det > kMaxBars)
        return perr(why, kCapacity, "detector bars %d outside [1, %d]", g.n_det, kMaxBars);
    if (g.nx < 1 || g.ny < 1 || g.nz < 1)
        return perr(why, kBadParam, "image dimensions %d x %d x %d", g.nx, g.ny, g.nz);
    g.n_vox = (int64_t)g.nx * g.ny * g.nz;
    g.wx = img[3];
    g.wy = img[4];
    g.wz = img[5];
    g.dphi = img[7];
    g.sx = img[8];
    g.sy = img[9];
    g.sz = img[10];
    g.fov2coll = img[11];
    const double F = g.fov2coll;

    hg.layers_d.resize(g.n_layers);
    hg.layers_f.resize(g.n_layers);
    int first = 0;
    for (int m = 0; m < g.n_layers; ++m) {
        const float* Lp = coll + (m + 1) * 10;
        int nh = (int)std::floor((double)Lp[0]);  // PESysMatGen.cu:982
        if (nh < 0) return perr(why, kBadParam, "layer %d: negative hole count", m);
        g.ranges[m].first = first;
        g.ranges[m].count = nh;
        first += nh;
        Layer<double>& L = hg.layers_d[m];
        L.lo[0] = -(double)Lp[1] / 2.0;
        L.hi[0] = (double)Lp[1] / 2.0;
        L.lo[1] = -(double)Lp[2] / 2.0 + F + (double)Lp[4];
        L.hi[1] = (double)Lp[2] / 2.0 + F + (double)Lp[4];
        L.lo[2] = -(double)Lp[3] / 2.0;
        L.hi[2] = (double)Lp[3] / 2.0;
        L.thick = Lp[2];
        L.mu = Lp[5];
        L.mupe = Lp[6];
        L.muc = Lp[7];
        Layer<float>& Lf = hg.layers_f[m];
        for (int k = 0; k < 3; ++k) {
            Lf.lo[k] = (float)L.lo[k];
            Lf.hi[k] = (float)L.hi[k];
        }
        Lf.thick = (float)L.thick;
        Lf.mu = (float)L.mu;
        Lf.mupe = (float)L.mupe;
        Lf.muc = (float)L.muc;
    }
    g.n_holes = first;
    if (g.n_holes > kMaxHoles)
        return perr(why, kCapacity, "%d holes exceed the parameter array (%d)", g.n_holes, kMaxHoles);
    hg.n_holes_l0 = (int)std::floor((double)coll[10] + 0.01);  // scatter.cu:2233
    if (g.n_layers == 0) hg.n_holes_l0 = 0;
    if (hg.n_holes_l0 > g.n_holes) hg.n_holes_l0 = g.n_holes;

    hg.holes_d.resize(g.n_holes);
    hg.holes_f.resize(g.n_holes);
    for (int h = 0; h < g.n_holes; ++h) {
        const float* hp = coll + 100 + 9 * h;
        Hole<double>& o = hg.holes_d[h];
        o.xc = hp[0];
        o.y1 = (double)hp[1] + F;
        o.y2 = (double)hp[2] + F;
        o.zc = hp[3];
        o.r = hp[4];
        o.mu = hp[5];
        o.mupe = hp[6];
        o.muc = hp[7];
        o.ymid = (o.y1 + o.y2) / 2.0;
        o.erx = o.erz = o.es = 0.0;
        o.ec = 1.0;
        hole_derived(hg, h);
    }

    hg.bars_d.resize(g.n_det);
    hg.bars_f.resize(g.n_det);
    for (int k = 0; k < g.n_det; ++k) {
        const float* b = det + 12 * k;
        Bar<double>& B = hg.bars_d[k];
        std::memset(&B, 0, sizeof B);
        B.cx = b[1];
        B.cy = (double)b[2] + F;
        B.cz = b[3];
        B.w = b[4];
        B.t = b[5];
        B.h = b[6];
        B.mut = b[7];
        B.mupe = b[8];
        B.muc = b[9];
        B.res = b[10];
        B.beta = b[11];
        B.cb = std::cos(B.beta);
        B.sb = std::sin(B.beta);
        B.rdiag = std::sqrt(B.w * B.w + B.h * B.h + B.t * B.t);
        B.pe_frac = B.mut != 0.0 ? B.mupe / B.mut : 0.0;
        B.pad1 = 0.0;
        if (!(B.w > 0 && B.t > 0 && B.h > 0))
            return perr(why, kBadParam, "bar %d has a non-positive size", k);
        Bar<float>& Bf = hg.bars_f[k];
        Bf.cx = (float)B.cx;
        Bf.cy = (float)B.cy;
        Bf.cz = (float)B.cz;
        Bf.w = (float)B.w;
        Bf.t = (float)B.t;
        Bf.h = (float)B.h;
        Bf.mut = (float)B.mut;
        Bf.mupe = (float)B.mupe;
        Bf.muc = (float)B.muc;
        Bf.res = (float)B.res;
        Bf.cb = (float)B.cb;
        Bf.sb = (float)B.sb;
        Bf.rdiag = (float)B.rdiag;
        Bf.beta = (float)B.beta;
        Bf.pe_frac = (float)B.pe_frac;
        Bf.pad1 = 0.0f;
    }
    make_group_boxes(hg);
    return 0;
}

// Records of the bitmaps and scatter kernels: bar bounding spheres, layer-0
// hole centres (scatter.cu:1465-1501, :1963-1991), scatter points.
inline void make_scatter_records(const HostGeom& hg, std::vector<Sphere<float>>& sf,
                          std::vector<Sphere<double>>& sd, std::vector<Sphere<float>>& hf,
                          std::vector<Sphere<double>>& hd, std::vector<ScatterPt<double>>& pb,
                          std::vector<ScatterPt<double>>& ph) {
    const GeomDesc& g = hg.g;
    const double F = g.fov2coll;
    for (int k = 0; k < g.n_det; ++k) {
        const float* b = hg.det.data() + 12 * k;
        double w = b[4], t = b[5], h = b[6];
        Sphere<double> S{(double)b[1], (double)b[2], (double)b[3], std::sqrt(w * w + h * h + t * t) / 2.0};
        sd.push_back(S);
        sf.push_back(Sphere<float>{(float)S.x, (float)S.y, (float)S.z, (float)S.r});
        const Bar<double>& B = hg.bars_d[k];
        pb.push_back(ScatterPt<double>{B.cx, B.cy, B.cz, 0.0});
    }
    for (int i = 0; i < hg.n_holes_l0; ++i) {
        const float* hp = hg.coll.data() + 100 + 9 * i;
        double y = ((double)hp[1] + (double)hp[2]) / 2.0;
        Sphere<double> S{(double)hp[0], y, (double)hp[3], 0.0};
        hd.push_back(S);
        hf.push_back(Sphere<float>{(float)S.x, (float)S.y, (float)S.z, 0.0f});
        ph.push_back(ScatterPt<double>{(double)hp[0], y + F, (double)hp[3],
                                       (double)hp[2] - (double)hp[1]});
    }
}

// Elliptical holes, an extension: the reference's elliptic-cylinder chord
// (PESysMatGen.cu:801-947) has no caller and the parameter file has no fields
// for it.  shapes holds x_rho, z_rho, theta (radians) per hole, n = the number
// of holes; a hole with x_rho <= 0 is the parameter file's circle; NULL makes
// every hole a circle again.  An elliptical hole's r is its bounding radius
// max(x_rho, z_rho): the culling predicate (its crit) and the candidate scans
// use it.  Returns 0, or kBadParam with the reason.
inline int apply_hole_shapes(HostGeom& hg, const float* shapes, int n, std::string& why) {
    const int nh = hg.g.n_holes;
    if (shapes && n != nh)
        return perr(why, kBadParam, "hole shapes given for %d holes, the geometry has %d", n, nh);
    for (int h = 0; h < nh; ++h) {
        const float* s = shapes ? shapes + 3 * h : nullptr;
        if (s && s[0] > 0.0f && !(s[1] > 0.0f && std::isfinite(s[0]) && std::isfinite(s[1]) &&
                                   std::isfinite(s[2])))
            return perr(why, kBadParam, "hole %d: semi-axes must be positive and finite, theta finite",
                        h);
    }
    for (int h = 0; h < nh; ++h) {
        Hole<double>& o = hg.holes_d[h];
        const float* s = shapes ? shapes + 3 * h : nullptr;
        if (s && s[0] > 0.0f) {
            o.erx = s[0];
            o.erz = s[1];
            o.ec = std::cos((double)s[2]);
            o.es = std::sin((double)s[2]);
            o.r = std::max(o.erx, o.erz);
        } else {
            o.erx = o.erz = o.es = 0.0;
            o.ec = 1.0;
            o.r = hg.coll[100 + 9 * h + 4];
        }
        hole_derived(hg, h);
    }
    make_group_boxes(hg);
    return 0;
}

}  // namespace smg
