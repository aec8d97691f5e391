"""ctypes binding of the FP64 host oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the CPU
baseline leg of the benchmark, where it is the checker.  The product path
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    V3 = C.POINTER(C.c_double)
    L.smg_o_geom_new.restype = C.c_void_p
    L.smg_o_geom_new.argtypes = [_fp, _fp, _fp, _fp]
    L.smg_o_geom_free.argtypes = [C.c_void_p]
    for name in ("smg_o_n_det", "smg_o_n_vox", "smg_o_n_holes_l0", "smg_o_n_layers"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.c_void_p]
    for name in ("smg_o_box_ray", "smg_o_box_ray_inside"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [_dp, _dp, _dp, _dp]
    L.smg_o_cylinder_ray.restype = C.c_double
    L.smg_o_cylinder_ray.argtypes = [_dp, _dp] + [C.c_double] * 5
    L.smg_o_ellcyl_ray.restype = C.c_double
    L.smg_o_ellcyl_ray.argtypes = [_dp, _dp] + [C.c_double] * 7
    L.smg_o_set_hole_shapes.restype = None
    L.smg_o_set_hole_shapes.argtypes = [C.c_void_p, C.c_void_p]
    L.smg_o_kn.restype = C.c_double
    L.smg_o_kn.argtypes = [C.c_double, C.c_double]
    L.smg_o_kn_sum.restype = C.c_double
    L.smg_o_kn_sum.argtypes = [C.c_double] * 4
    L.smg_o_cone_angle.restype = C.c_double
    L.smg_o_cone_angle.argtypes = [_dp, _dp, _dp]
    L.smg_o_scatter_energy.restype = C.c_double
    L.smg_o_scatter_energy.argtypes = [C.c_double, C.c_double]
    L.smg_o_gauss_window.restype = C.c_double
    L.smg_o_gauss_window.argtypes = [C.c_double] * 4
    L.smg_o_pe_element.restype = C.c_double
    L.smg_o_pe_element.argtypes = [C.c_void_p] + [C.c_int] * 6
    L.smg_o_pe_rows.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                _dp, C.c_int]
    L.smg_o_pe_pairs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip,
                                 C.c_int, _dp, C.c_int]
    L.smg_o_scatter_ops_per_term.restype = C.c_double
    L.smg_o_scatter_ops_per_term.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_uint,
                                             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, _ip, _ip, C.c_int, C.POINTER(C.c_double)]
    L.smg_o_pe_ops_per_ray.restype = C.c_double
    L.smg_o_pe_ops_per_ray.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip,
                                       C.c_int]
    L.smg_o_pe_cull_counts.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.smg_o_pe_grad.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                C.POINTER(C.c_double), C.c_int]
    L.smg_o_bitmap_crystal_words.restype = C.c_int64
    L.smg_o_bitmap_crystal_words.argtypes = [C.c_int]
    L.smg_o_bitmap_collimator_words.restype = C.c_int64
    L.smg_o_bitmap_collimator_words.argtypes = [C.c_int, C.c_int]
    L.smg_o_bitmap_crystal.argtypes = [C.c_void_p, _up, C.c_int]
    L.smg_o_bitmap_collimator.argtypes = [C.c_void_p, _up, C.c_int]
    L.smg_o_crystal_pred.restype = C.c_int
    L.smg_o_crystal_pred.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.smg_o_coll_pred.restype = C.c_int
    L.smg_o_coll_pred.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.smg_o_bit_crystal.restype = C.c_int
    L.smg_o_bit_crystal.argtypes = [_up, C.c_int, C.c_int, C.c_int, C.c_int]
    L.smg_o_bit_collimator.restype = C.c_int
    L.smg_o_bit_collimator.argtypes = [_up, C.c_int, C.c_int, C.c_int, C.c_int]
    L.smg_o_crystal_scatter_element.restype = C.c_double
    L.smg_o_seg_box.restype = C.c_double
    L.smg_o_seg_box.argtypes = [_dp, _dp, _dp, _dp]
    L.smg_o_seg_cyl.restype = C.c_double
    L.smg_o_seg_cyl.argtypes = [_dp, _dp] + [C.c_double] * 5
    L.smg_o_seg_ell.restype = C.c_double
    L.smg_o_seg_ell.argtypes = [_dp, _dp] + [C.c_double] * 7
    L.smg_o_coll_grid_points.restype = C.c_int64
    L.smg_o_coll_grid_points.argtypes = [C.c_void_p, C.c_double, C.c_uint, C.c_void_p, C.c_int64]
    L.smg_o_coll_grid_weight.restype = C.c_double
    L.smg_o_coll_grid_weight.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    L.smg_o_coll_grid_scatter_pairs.argtypes = [C.c_void_p, C.c_double, C.c_uint, C.c_int, C.c_int,
                                                C.c_int, C.c_int, _ip, _ip, C.c_int, _dp, C.c_int]
    L.smg_o_crystal_grid_weight.restype = C.c_double
    L.smg_o_crystal_grid_weight.argtypes = [C.c_void_p] + [C.c_int] * 7
    L.smg_o_crystal_grid_scatter_pairs.argtypes = [C.c_void_p, _up] + [C.c_int] * 7 + [
        _ip, _ip, C.c_int, _dp, C.c_int]
    L.smg_o_set_mu_tables.restype = C.c_int
    L.smg_o_set_mu_tables.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.smg_o_mu_ratio.restype = C.c_double
    L.smg_o_mu_ratio.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    L.smg_o_crystal_scatter_element.argtypes = [C.c_void_p, _fp, _up] + [C.c_int] * 6
    L.smg_o_collimator_scatter_element.restype = C.c_double
    L.smg_o_collimator_scatter_element.argtypes = [C.c_void_p, _up] + [C.c_int] * 7
    L.smg_o_crystal_scatter_pairs.argtypes = [C.c_void_p, _fp, _up, C.c_int, C.c_int, C.c_int,
                                              C.c_int, _ip, _ip, C.c_int, _dp, C.c_int]
    L.smg_o_collimator_scatter_pairs.argtypes = [C.c_void_p, _up, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, _ip, _ip, C.c_int, _dp,
                                                 C.c_int]
    L.smg_o_max_threads.restype = C.c_int
    _lib = L
    return L


def _v3(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))


# ---- primitives -------------------------------------------------------------------

def box_ray(a, b, lo, hi) -> float:
    return lib().smg_o_box_ray(_v3(a), _v3(b), _v3(lo), _v3(hi))


def box_ray_inside(a, b, lo, hi) -> float:
    return lib().smg_o_box_ray_inside(_v3(a), _v3(b), _v3(lo), _v3(hi))


def cylinder_ray(a, b, xc, y1, y2, zc, r) -> float:
    return lib().smg_o_cylinder_ray(_v3(a), _v3(b), xc, y1, y2, zc, r)


def ellcyl_ray(a, b, xc, y1, y2, zc, x_rho, z_rho, theta) -> float:
    return lib().smg_o_ellcyl_ray(_v3(a), _v3(b), xc, y1, y2, zc, x_rho, z_rho, theta)


def seg_box(a, b, lo, hi) -> float:
    """Length of segment a -> b inside the box (no reference thresholds)."""
    return lib().smg_o_seg_box(_v3(a), _v3(b), _v3(lo), _v3(hi))


def seg_cyl(a, b, xc, y1, y2, zc, r) -> float:
    """Length of segment a -> b inside the finite cylinder along y."""
    return lib().smg_o_seg_cyl(_v3(a), _v3(b), xc, y1, y2, zc, r)


def seg_ell(a, b, xc, y1, y2, zc, x_rho, z_rho, theta) -> float:
    """Length of segment a -> b inside the finite elliptic cylinder along y."""
    return lib().smg_o_seg_ell(_v3(a), _v3(b), xc, y1, y2, zc, x_rho, z_rho, theta)


def kn(theta, e0) -> float:
    return lib().smg_o_kn(theta, e0)


def kn_sum(e0, lo, hi, step) -> float:
    return lib().smg_o_kn_sum(e0, lo, hi, step)


def cone_angle(src, a, det) -> float:
    return lib().smg_o_cone_angle(_v3(src), _v3(a), _v3(det))


def scatter_energy(theta, e0) -> float:
    return lib().smg_o_scatter_energy(theta, e0)


def gauss_window(e, res, lo, hi) -> float:
    return lib().smg_o_gauss_window(e, res, lo, hi)


def max_threads() -> int:
    return lib().smg_o_max_threads()


# ---- geometry-level oracle ---------------------------------------------------------

class Oracle:
    """FP64 oracle bound to one set of parameter arrays."""

    def __init__(self, params):
        L = lib()
        self._arrays = [np.ascontiguousarray(a, dtype=np.float32)
                        for a in (params.coll, params.det, params.img, params.phys)]
        self._g = L.smg_o_geom_new(*self._arrays)
        if not self._g:
            raise ValueError("oracle rejected the parameter arrays")
        self.n_det = L.smg_o_n_det(self._g)
        self.n_vox = L.smg_o_n_vox(self._g)
        self.n_holes_l0 = L.smg_o_n_holes_l0(self._g)
        self.n_layers = L.smg_o_n_layers(self._g)

    def __del__(self):
        g = getattr(self, "_g", None)
        if g:
            lib().smg_o_geom_free(g)
            self._g = None

    def set_hole_shapes(self, shapes=None):
        """Elliptical holes: (x_rho, z_rho, theta) per hole, x_rho <= 0 keeps a
        circle; None: all circles."""
        if shapes is None:
            self._shapes = None
            lib().smg_o_set_hole_shapes(self._g, None)
            return
        self._shapes = np.ascontiguousarray(shapes, np.float32).reshape(-1)
        lib().smg_o_set_hole_shapes(self._g, self._shapes.ctypes.data)

    def pe_element(self, d, v, rot=0, sub=(8, 16, 8)) -> float:
        return lib().smg_o_pe_element(self._g, d, v, rot, *sub)

    def pe_rows(self, rows, rot=0, sub=(8, 16, 8), nthreads=0) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        out = np.zeros((rows.size, self.n_vox), np.float64)
        lib().smg_o_pe_rows(self._g, rot, *sub, rows, rows.size, out, nthreads)
        return out

    def pe_matrix(self, rot=0, sub=(8, 16, 8), nthreads=0) -> np.ndarray:
        return self.pe_rows(np.arange(self.n_det), rot, sub, nthreads)

    def pe_pairs(self, d, v, rot=0, sub=(8, 16, 8), nthreads=0) -> np.ndarray:
        d = np.ascontiguousarray(d, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        out = np.zeros(d.size, np.float64)
        lib().smg_o_pe_pairs(self._g, rot, *sub, d, v, d.size, out, nthreads)
        return out

    def pe_ops_per_ray(self, d, v, rot=0, sub=(8, 16, 8)) -> float:
        d = np.ascontiguousarray(d, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        return lib().smg_o_pe_ops_per_ray(self._g, rot, *sub, d, v, d.size)

    def scatter_ops_per_term(self, kind, d, v, bitmap=None, pitch=0.5, layers=1, grid=(1, 1, 1),
                             rot=0, sub=(1, 1, 1)):
        """(algorithmic ops per scatter term, terms counted) of the reference's
        scatter element body for rows d and voxels v; kind 0 crystal, 1
        collimator (hole centres), 2 collimator grid, 3 crystal grid."""
        d = np.ascontiguousarray(d, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        bm = None if bitmap is None else np.ascontiguousarray(bitmap, np.uint32)
        terms = C.c_double()
        ops = lib().smg_o_scatter_ops_per_term(
            self._g, None if bm is None else bm.ctypes.data_as(C.c_void_p), kind, pitch, layers,
            *grid, rot, *sub, d, v, d.size, C.byref(terms))
        return ops, terms.value

    def cull_counts(self, d, v, rot=0):
        nc, nh = C.c_int(), C.c_int()
        lib().smg_o_pe_cull_counts(self._g, d, v, rot, C.byref(nc), C.byref(nh))
        return nc.value, nh.value

    def pe_grad(self, w, rot=0, sub=(8, 16, 8), nthreads=0):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.n_det, self.n_vox)
        grad = np.zeros(self.n_layers + 2 * self.n_det, np.float64)
        loss = C.c_double()
        lib().smg_o_pe_grad(self._g, rot, *sub, w, grad, C.byref(loss), nthreads)
        return loss.value, grad

    def bitmap_crystal(self, nthreads=0) -> np.ndarray:
        L = lib()
        words = np.zeros(L.smg_o_bitmap_crystal_words(self.n_det), np.uint32)
        L.smg_o_bitmap_crystal(self._g, words, nthreads)
        return words

    def bitmap_collimator(self, nthreads=0) -> np.ndarray:
        L = lib()
        words = np.zeros(L.smg_o_bitmap_collimator_words(self.n_holes_l0, self.n_det), np.uint32)
        L.smg_o_bitmap_collimator(self._g, words, nthreads)
        return words

    def crystal_pred(self, i, j, k) -> int:
        return lib().smg_o_crystal_pred(self._g, i, j, k)

    def coll_pred(self, h, j, k) -> int:
        return lib().smg_o_coll_pred(self._g, h, j, k)

    def set_mu_tables(self, energy=None, crystal_pe=None, crystal_c=None, coll_pe=None,
                      coll_c=None):
        """mu(E) tables (keV, increasing; one array per process); None clears them."""
        if energy is None:
            lib().smg_o_set_mu_tables(self._g, 0, *([np.zeros(1)] * 5))
            return
        arrs = [np.ascontiguousarray(a, np.float64) for a in
                (energy, crystal_pe, crystal_c, coll_pe, coll_c)]
        if lib().smg_o_set_mu_tables(self._g, len(arrs[0]), *arrs) != 0:
            raise ValueError("bad mu(E) tables")

    def mu_ratio(self, which, e, e0) -> float:
        return lib().smg_o_mu_ratio(self._g, which, e, e0)

    def coll_grid_points(self, pitch, layers=1):
        """(points [n][3], weights [n]) of the grid over the plates of the
        layers in the bit mask `layers`; weight = cell volume x the layer's
        Compton mu."""
        n = lib().smg_o_coll_grid_points(self._g, pitch, layers, None, 0)
        pts = np.zeros((max(n, 1), 4))
        lib().smg_o_coll_grid_points(self._g, pitch, layers, pts.ctypes.data_as(C.c_void_p), n)
        return pts[:n, :3], pts[:n, 3]

    def coll_grid_weight(self, p, s, w) -> float:
        return lib().smg_o_coll_grid_weight(self._g, _v3(p), _v3(s), w)

    def coll_grid_scatter_pairs(self, pitch, d, v, rot=0, sub=(1, 1, 1), nthreads=0, layers=1):
        d = np.ascontiguousarray(d, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        out = np.zeros(len(d))
        lib().smg_o_coll_grid_scatter_pairs(self._g, pitch, layers, rot, *sub, d, v, len(d), out,
                                            nthreads)
        return out

    def crystal_grid_weight(self, b, v, grid, cell, rot=0) -> float:
        return lib().smg_o_crystal_grid_weight(self._g, b, v, rot, *grid, cell)

    def crystal_grid_scatter_pairs(self, bitmap, grid, d, v, rot=0, sub=(1, 1, 1), nthreads=0):
        d = np.ascontiguousarray(d, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        out = np.zeros(len(d))
        bm = np.ascontiguousarray(bitmap, np.uint32)
        lib().smg_o_crystal_grid_scatter_pairs(self._g, bm, rot, *grid, *sub, d, v, len(d), out,
                                                nthreads)
        return out

    def crystal_scatter_pairs(self, pe, bitmap, d, v, rot=0, sub=(1, 1, 1), nthreads=0):
        pe = np.ascontiguousarray(pe, dtype=np.float32).reshape(-1)
        assert pe.size >= self.n_det * self.n_vox
        bitmap = np.ascontiguousarray(bitmap, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        out = np.zeros(d.size, np.float64)
        lib().smg_o_crystal_scatter_pairs(self._g, pe, bitmap, rot, *sub, d, v, d.size, out,
                                          nthreads)
        return out

    def collimator_scatter_pairs(self, bitmap, d, v, rot=0, variant=0, sub=(1, 1, 1),
                                 nthreads=0):
        bitmap = np.ascontiguousarray(bitmap, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        out = np.zeros(d.size, np.float64)
        lib().smg_o_collimator_scatter_pairs(self._g, bitmap, rot, variant, *sub, d, v, d.size,
                                             out, nthreads)
        return out
