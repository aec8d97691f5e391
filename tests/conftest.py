"""Shared test setup.

`-m "not gpu"` tests run anywhere (oracle against closed forms and golden
vectors, host logic, the C ABI's exports).  `-m gpu` tests call the HIP
kernels through the C ABI on an MI355X and compare with the FP64 oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-based-system-matrix-calculation-for-spect-pet_amd")
ORACLE = os.path.join(ROOT, "oracle")
for path in (PKG, ORACLE):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP kernels through the C ABI)")


# Relative-error bar of the FP32 path against the FP64 oracle (the benchmark
# metric's <= 1e-5 max relative error), and of the FP64 path (output rounded
# to float32: half an ulp plus the oracle's own rounding).
TOL_FP32 = 1e-5
TOL_FP64 = 2e-7
TINY = 1e-30


def rel_errors(got, ref):
    """Max and median relative error over |ref| >= 1e-30; elements below must
    match to 1e-30 absolute."""
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    big = np.abs(ref) >= TINY
    small_ok = bool(np.all(np.abs(got[~big] - ref[~big]) <= TINY))
    if not big.any():
        return 0.0, 0.0, small_ok
    rel = np.abs(got[big] - ref[big]) / np.abs(ref[big])
    return float(rel.max()), float(np.median(rel)), small_ok


def assert_parity(got, ref, tol, what=""):
    mx, med, small_ok = rel_errors(got, ref)
    print(f"{what}: max rel {mx:.3e} median {med:.3e}", flush=True)
    assert small_ok, f"{what}: elements with |ref| < 1e-30 differ"
    assert mx <= tol, f"{what}: max relative error {mx:.3e} > {tol:.1e} (median {med:.3e})"


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def params_mod():
    from smg import params
    return params


def elliptic_hole_shapes(n_holes, seed=0, circles_every=4):
    """(x_rho, z_rho, theta) per hole: ellipses of semi-axes 0.35-1.0 mm at
    random turns, every circles_every-th hole left circular (x_rho = 0)."""
    rng = np.random.default_rng(seed)
    s = np.zeros((n_holes, 3), np.float32)
    s[:, 0] = rng.uniform(0.5, 1.0, n_holes)
    s[:, 1] = rng.uniform(0.35, 0.7, n_holes)
    s[:, 2] = rng.uniform(-np.pi, np.pi, n_holes)
    if circles_every:
        s[::circles_every, 0] = 0.0
    return s


@pytest.fixture(scope="session")
def lane_lib():
    """Host build of the GPU lane arithmetic (tests/lane_harness.cpp)."""
    return load_lane_lib()


def load_lane_lib():
    """Builds (when a source is newer) and loads tests/lane_harness.cpp."""
    import ctypes as C
    out_dir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out_dir, exist_ok=True)
    so = os.path.join(out_dir, "liblane.so")
    src = os.path.join(ROOT, "tests", "lane_harness.cpp")
    hdrs = [os.path.join(PKG, "csrc", f) for f in os.listdir(os.path.join(PKG, "csrc"))
            if f.endswith(".hpp")]
    newest = max(os.path.getmtime(p) for p in hdrs + [src])
    if not os.path.exists(so) or os.path.getmtime(so) < newest:
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas",
                        "-I", os.path.join(PKG, "csrc"), "-o", so, src], check=True)
    L = C.CDLL(so)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    up = np.ctypeslib.ndpointer(np.uint32, flags="C")
    L.lane_seg_ell.restype = C.c_double
    L.lane_seg_ell.argtypes = [dp, dp] + [C.c_double] * 7
    L.lane_xcd_items.restype = None
    L.lane_xcd_items.argtypes = [C.c_int, ip]
    L.lane_pe_pairs.restype = C.c_long
    L.lane_pe_pairs.argtypes = [fp, fp, fp, fp] + [C.c_int] * 5 + [ip, ip, C.c_int, dp]
    L.lane_narrow_check.restype = C.c_int
    L.lane_narrow_check.argtypes = [fp, fp, fp, fp] + [C.c_int] * 8 + [C.c_uint, dp]
    L.lane_kn_sum_table.restype = None
    L.lane_kn_sum_table.argtypes = [C.c_double, fp, fp, C.c_int, dp, ip]
    L.lane_scatter_pairs.restype = C.c_long
    L.lane_scatter_pairs.argtypes = [fp, fp, fp, fp, fp, up] + [C.c_int] * 7 + [ip, ip, C.c_int, dp]
    L.lane_bitmap.restype = C.c_long
    L.lane_mu_ratio_f.restype = C.c_int
    L.lane_mu_ratio_f.argtypes = [C.c_double, C.c_int] + [dp] * 6 + [C.c_int, dp, dp]
    L.lane_window_reject_cos.restype = C.c_double
    L.lane_window_reject_cos.argtypes = [C.c_double] * 4
    L.lane_window_cut_margin.restype = C.c_double
    L.lane_window_cut_margin.argtypes = [C.c_double] * 2
    L.lane_set_hole_shapes.restype = None
    L.lane_set_hole_shapes.argtypes = [fp, C.c_int]
    L.lane_bitmap.argtypes = [fp, fp, fp, fp, C.c_int, up, C.c_long]
    return L


@pytest.fixture(scope="session")
def smg_mod():
    """The product library; gpu tests fail loudly if it is not built."""
    import smg
    smg.lib()
    return smg


def small_geometry(params_mod, beta=0.0, layers=1, nx=13, ny=11, nz=3, dphi=0.0, n_rot=1):
    """A small geometry exercising rotated bars, several collimator layers and
    ragged voxel tiles (13 x 11 x 3 does not divide into 64-voxel tiles)."""
    p = params_mod.Params.empty()
    p.coll[0] = layers
    rng = np.random.default_rng(11)
    h = 0
    for m in range(layers):
        nh = 9
        params_mod.set_layer(p, m, nh, 40.0, 3.0, 40.0, 6.0 * m, 3.65, 3.3, 0.35)
        for i in range(nh):
            x, z = rng.uniform(-6, 6, 2)
            params_mod.set_hole(p, h, x, 6.0 * m - 1.5, 6.0 * m + 1.5, z, rng.uniform(0.4, 0.9))
            h += 1
    p.det[0] = 18
    for k in range(18):
        ix, iz = k % 6, k // 6
        params_mod.set_bar(p, k, -5.0 + 2.0 * ix, 12.0 + 6.0 * (layers - 1) + (k % 2) * 0.5,
                           -2.0 + 2.0 * iz,
                           1.8, 6.0, 1.9, 0.25, 0.18, 0.07, 0.1, beta * (1 + 0.1 * k))
    params_mod.set_image(p, nx, ny, nz, 0.8, 0.9, 1.1, n_rot=n_rot, dphi=dphi,
                         shift=(0.3, -0.2, 0.1), fov2coll=25.0)
    params_mod.set_physics(p)
    return p


def narrow_plate_geometry(params_mod):
    """small_geometry with two plates of 9 x 7 mm (x by z): the plate edges cut
    through the bundle of rays, so some rays miss a plate or leave through its
    side and are stopped by the plate gate."""
    p = small_geometry(params_mod, beta=0.05, layers=2)
    for m in range(2):
        p.coll[(m + 1) * 10 + 1] = 9.0
        p.coll[(m + 1) * 10 + 3] = 7.0
    return p


def overflow_geometry(params_mod, n_layers=24, bars_per_side=12):
    """A geometry whose voxel tiles need more candidates than the PE kernels'
    LDS lists hold (448 bars, 256 holes; 320 bars in the adjoint): a 20 mm tungsten plate with 1,600
    holes on a 1 mm grid, a deep block of n_layers layers of 2 mm cubic bars
    right behind it, and 8 mm voxels, so a 4x4x4-voxel tile sees a wide
    frustum through a thick plate.  The kernels then redo such tiles with
    candidate lists in global memory."""
    p = params_mod.Params.empty()
    p.coll[0] = 1
    T, nh = 20.0, 40
    params_mod.set_layer(p, 0, nh * nh, 80.0, T, 80.0, 0.0, **params_mod.TUNGSTEN)
    h = 0
    for iz in range(nh):
        for ix in range(nh):
            params_mod.set_hole(p, h, (ix - (nh - 1) / 2) * 1.0, -T / 2, T / 2,
                                (iz - (nh - 1) / 2) * 1.0, 0.3)
            h += 1
    n = bars_per_side
    p.det[0] = n_layers * n * n
    k = 0
    for layer in range(n_layers):
        for iz in range(n):
            for ix in range(n):
                params_mod.set_bar(p, k, (ix - (n - 1) / 2) * 2.0, T / 2 + 6.0 + 2.0 * layer,
                                   (iz - (n - 1) / 2) * 2.0, 2.0, 2.0, 2.0, res=0.1, **params_mod.BAR_MU)
                k += 1
    params_mod.set_image(p, 8, 8, 4, 8.0, 8.0, 8.0, fov2coll=40.0)
    params_mod.set_physics(p)
    return p
