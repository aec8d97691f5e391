"""Python binding of libsmg, the C ABI of the MI355X system-matrix generator.

The host-side mirror of the reference's launch wrappers
(PEGen_RayTracing_CircularHole/PESysMatGen.cu:1368-1437 and the three
``scatter()`` functions): a :class:`Context` owns one GPU and the parsed
geometry; its methods return numpy arrays on the host or write into device
buffers (``torch.Tensor.data_ptr()``) directly.

Loading fails loudly when ``libsmg.so`` has not been built; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import params  # noqa: F401  (re-exported)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsmg.so")

SMG_OK = 0
STATUS_NAMES = {0: "OK", 1: "BAD_PARAM", 2: "CAPACITY", 3: "IO", 4: "HIP", 5: "NO_DEVICE",
                6: "STATE"}
PREC = {"fp32": 0, "fp64": 1}
COLL_VARIANT = {"combined": 0, "standalone": 1}

# Every entry point declared in include/smg.h.
EXPORTS = (
    "smg_device_count", "smg_create", "smg_destroy", "smg_last_error", "smg_status_string",
    "smg_stream", "smg_synchronize", "smg_set_geometry", "smg_get_dims",
    "smg_set_pe_subsamples", "smg_set_scatter_subsamples", "smg_pe_forward", "smg_pe_backward",
    "smg_pe_project", "smg_pe_backproject",
    "smg_bitmap_crystal_words", "smg_bitmap_collimator_words", "smg_bitmap_crystal",
    "smg_bitmap_collimator", "smg_load_bitmap_crystal", "smg_load_bitmap_collimator",
    "smg_scatter_crystal", "smg_scatter_collimator", "smg_set_mu_tables",
    "smg_scatter_collimator_grid", "smg_scatter_crystal_grid", "smg_set_hole_shapes",
    "smg_pe_forward_noalloc", "smg_scatter_sum", "smg_pe_tangent",
)


class SmgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"smg status {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("elements", C.c_uint64),
                ("rays", C.c_uint64), ("fp64_fallbacks", C.c_uint64),
                ("overflow_waves", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib():
    """Load libsmg.so (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C {PKG_DIR}` "
                          "(or __graft_entry__.build())")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (same
    # soname, libamdhip64.so.7) and loads it by an unversioned name, so if
    # libsmg pulled in the system copy first, torch would load a second
    # runtime that cannot open the GPU.  Importing torch first makes libsmg
    # bind to torch's copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, sz = C.c_void_p, C.c_int, C.c_int64, C.c_size_t
    fp, dp, up = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_uint32)
    st = C.POINTER(Stats)
    sig = {
        "smg_device_count": (i32, [C.POINTER(C.c_int)]),
        "smg_create": (i32, [i32, C.POINTER(vp)]),
        "smg_destroy": (i32, [vp]),
        "smg_last_error": (C.c_char_p, [vp]),
        "smg_status_string": (C.c_char_p, [i32]),
        "smg_stream": (vp, [vp]),
        "smg_synchronize": (i32, [vp]),
        "smg_set_geometry": (i32, [vp, fp, sz, fp, sz, fp, sz, fp, sz]),
        "smg_get_dims": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
        "smg_set_pe_subsamples": (i32, [vp, i32, i32, i32]),
        "smg_set_scatter_subsamples": (i32, [vp, i32, i32, i32]),
        "smg_set_mu_tables": (i32, [vp, i32, dp, dp, dp, dp, dp]),
        "smg_set_hole_shapes": (i32, [vp, fp, i32]),
        "smg_pe_forward": (i32, [vp, i32, i32, vp, i32, st]),
        "smg_pe_forward_noalloc": (i32, [vp, i32, i32, vp, st]),
        "smg_pe_backward": (i32, [vp, i32, i32, vp, i32, dp, dp, st]),
        "smg_pe_tangent": (i32, [vp, i32, i32, dp, vp, vp, i32, st]),
        "smg_pe_project": (i32, [vp, i32, i32, vp, i32, vp, i32, st]),
        "smg_pe_backproject": (i32, [vp, i32, i32, vp, i32, vp, i32, st]),
        "smg_bitmap_crystal_words": (i64, [i32]),
        "smg_bitmap_collimator_words": (i64, [i32, i32]),
        "smg_bitmap_crystal": (i32, [vp, vp, i64, i32]),
        "smg_bitmap_collimator": (i32, [vp, vp, i64, i32]),
        "smg_load_bitmap_crystal": (i32, [vp, up, i64]),
        "smg_load_bitmap_collimator": (i32, [vp, up, i64]),
        "smg_scatter_crystal": (i32, [vp, i32, i32, vp, i32, vp, i32, st]),
        "smg_scatter_collimator": (i32, [vp, i32, i32, i32, vp, i32, st]),
        "smg_scatter_sum": (i32, [vp, i32, C.c_uint, i32, vp, vp, i32, vp, vp, i32, st]),
        "smg_scatter_collimator_grid": (i32, [vp, i32, C.c_double, C.c_uint, i32, vp, i32,
                                              C.POINTER(C.c_int64), st]),
        "smg_scatter_crystal_grid": (i32, [vp, i32, i32, i32, i32, vp, i32, i32, vp, i32, st]),
    }
    missing = [name for name in sig if not hasattr(L, name)]
    if missing:
        raise ImportError(f"{LIB_PATH} lacks entry points {missing}: rebuild it")
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def device_count() -> int:
    n = C.c_int(0)
    lib().smg_device_count(C.byref(n))
    return n.value


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class Dims:
    n_det: int
    n_vox: int
    n_rot: int
    n_holes_l0: int


class Context:
    """One GPU, one stream, one geometry."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        s = self._L.smg_create(device, C.byref(h))
        if s != SMG_OK:
            raise SmgError(s, f"smg_create(device={device}) failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.smg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status: int):
        if status != SMG_OK:
            raise SmgError(status, self._L.smg_last_error(self._h).decode())

    @property
    def stream(self) -> int:
        return self._L.smg_stream(self._h) or 0

    def synchronize(self):
        self._check(self._L.smg_synchronize(self._h))

    # ---- geometry ----------------------------------------------------------------
    def set_geometry(self, p) -> Dims:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (p.coll, p.det, p.img, p.phys)]
        self._check(self._L.smg_set_geometry(self._h, _fptr(arrs[0]), arrs[0].size,
                                             _fptr(arrs[1]), arrs[1].size, _fptr(arrs[2]),
                                             arrs[2].size, _fptr(arrs[3]), arrs[3].size))
        self.dims = self.get_dims()
        self._n_layers = p.n_layers
        return self.dims

    def get_dims(self) -> Dims:
        nd, nv, nr, nh = C.c_int(), C.c_int64(), C.c_int(), C.c_int()
        self._check(self._L.smg_get_dims(self._h, C.byref(nd), C.byref(nv), C.byref(nr),
                                         C.byref(nh)))
        return Dims(nd.value, nv.value, nr.value, nh.value)

    def set_pe_subsamples(self, sx: int, sy: int, sz: int):
        self._check(self._L.smg_set_pe_subsamples(self._h, sx, sy, sz))

    def set_scatter_subsamples(self, sx: int, sy: int, sz: int):
        self._check(self._L.smg_set_scatter_subsamples(self._h, sx, sy, sz))

    def set_mu_tables(self, energy=None, crystal_pe=None, crystal_compton=None, coll_pe=None,
                      coll_compton=None):
        """mu(E) tables (node energies in keV, increasing; one value array per
        process) for the scatter kernels; None restores the power laws."""
        if energy is None:
            z = np.zeros(1)
            self._check(self._L.smg_set_mu_tables(self._h, 0, *[_dptr(z)] * 5))
            return
        arrs = [np.ascontiguousarray(a, np.float64) for a in
                (energy, crystal_pe, crystal_compton, coll_pe, coll_compton)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise ValueError("mu(E) tables must have one value per node")
        self._check(self._L.smg_set_mu_tables(self._h, n, *[_dptr(a) for a in arrs]))

    def set_hole_shapes(self, shapes=None):
        """Elliptical holes: shapes[h] = (x_rho, z_rho, theta) for every hole of
        the geometry (x_rho <= 0 keeps hole h circular); None makes every hole
        a circle.  A new geometry resets the shapes."""
        if shapes is None:
            self._check(self._L.smg_set_hole_shapes(self._h, None, 0))
            return
        s = np.ascontiguousarray(shapes, np.float32).reshape(-1, 3)
        self._check(self._L.smg_set_hole_shapes(
            self._h, s.ctypes.data_as(C.POINTER(C.c_float)), s.shape[0]))

    # ---- PE ----------------------------------------------------------------------
    def pe_forward(self, rotation: int = 0, precision: str = "fp32", out=None,
                   device_ptr: int | None = None) -> tuple[np.ndarray | None, dict]:
        """PE block of one rotation.  With ``device_ptr`` (a device buffer of
        N*V floats) the result stays on the GPU and None is returned."""
        st = Stats()
        if device_ptr is not None:
            self._check(self._L.smg_pe_forward(self._h, rotation, PREC[precision],
                                               C.c_void_p(device_ptr), 1, C.byref(st)))
            return None, st.as_dict()
        n = self.dims.n_det * self.dims.n_vox
        if out is None:
            out = np.empty((self.dims.n_det, self.dims.n_vox), np.float32)
        assert out.dtype == np.float32 and out.size == n and out.flags.c_contiguous
        self._check(self._L.smg_pe_forward(self._h, rotation, PREC[precision],
                                           out.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
        return out, st.as_dict()

    def pe_backward(self, w: np.ndarray, rotation: int = 0,
                    precision: str = "fp32") -> tuple[float, np.ndarray, dict]:
        """Loss sum(w * A) and its gradient: [layer mu][bar mu_total][bar mu_pe]."""
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(self.dims.n_det, self.dims.n_vox)
        n_par = self._n_layers + 2 * self.dims.n_det
        grad = np.zeros(n_par, np.float64)
        loss = C.c_double()
        st = Stats()
        self._check(self._L.smg_pe_backward(self._h, rotation, PREC[precision],
                                            w.ctypes.data_as(C.c_void_p), 0,
                                            grad.ctypes.data_as(C.POINTER(C.c_double)),
                                            C.byref(loss), C.byref(st)))
        return loss.value, grad, st.as_dict()

    def pe_tangent(self, dmu: np.ndarray, rotation: int = 0, precision: str = "fp32",
                   with_matrix: bool = False):
        """dA = sum_k dmu[k] dA/dmu_k (dmu in gradient order); returns (dA, A or
        None, stats)."""
        n_par = self._n_layers + 2 * self.dims.n_det
        dmu = np.ascontiguousarray(dmu, np.float64).reshape(n_par)
        shape = (self.dims.n_det, self.dims.n_vox)
        dA = np.empty(shape, np.float32)
        A = np.empty(shape, np.float32) if with_matrix else None
        st = Stats()
        self._check(self._L.smg_pe_tangent(self._h, rotation, PREC[precision],
                                           dmu.ctypes.data_as(C.POINTER(C.c_double)),
                                           dA.ctypes.data_as(C.c_void_p),
                                           None if A is None else A.ctypes.data_as(C.c_void_p), 0,
                                           C.byref(st)))
        return dA, A, st.as_dict()

    def pe_project(self, x: np.ndarray, rotation: int = 0,
                   precision: str = "fp32") -> tuple[np.ndarray, dict]:
        """y = A x with the PE matrix of `rotation`, matrix-free (x: V values)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.dims.n_vox)
        y = np.empty(self.dims.n_det, np.float32)
        st = Stats()
        self._check(self._L.smg_pe_project(self._h, rotation, PREC[precision],
                                           x.ctypes.data_as(C.c_void_p), 0,
                                           y.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
        return y, st.as_dict()

    def pe_backproject(self, y: np.ndarray, rotation: int = 0,
                       precision: str = "fp32") -> tuple[np.ndarray, dict]:
        """x = A^T y with the PE matrix of `rotation`, matrix-free (y: N values)."""
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(self.dims.n_det)
        x = np.empty(self.dims.n_vox, np.float32)
        st = Stats()
        self._check(self._L.smg_pe_backproject(self._h, rotation, PREC[precision],
                                               y.ctypes.data_as(C.c_void_p), 0,
                                               x.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
        return x, st.as_dict()

    # ---- bitmaps -----------------------------------------------------------------
    def bitmap_crystal(self, device_ptr: int | None = None) -> np.ndarray | None:
        """Compute the crystal bitmap; copy it to a new host array, or to the
        device buffer at device_ptr (then None is returned)."""
        n = self._L.smg_bitmap_crystal_words(self.dims.n_det)
        if device_ptr is not None:
            self._check(self._L.smg_bitmap_crystal(self._h, C.c_void_p(device_ptr), n, 1))
            return None
        words = np.zeros(n, np.uint32)
        self._check(self._L.smg_bitmap_crystal(self._h, words.ctypes.data_as(C.c_void_p), n, 0))
        return words

    def bitmap_collimator(self, device_ptr: int | None = None) -> np.ndarray | None:
        n = self._L.smg_bitmap_collimator_words(self.dims.n_holes_l0, self.dims.n_det)
        if device_ptr is not None:
            self._check(self._L.smg_bitmap_collimator(self._h, C.c_void_p(device_ptr), n, 1))
            return None
        words = np.zeros(n, np.uint32)
        self._check(self._L.smg_bitmap_collimator(self._h, words.ctypes.data_as(C.c_void_p), n, 0))
        return words

    def build_bitmap(self, collimator: bool = False) -> float:
        """Compute a bitmap into the context only (no copy out); returns the
        wall time of the call in ms (kernel plus synchronisation)."""
        import time
        fn = self._L.smg_bitmap_collimator if collimator else self._L.smg_bitmap_crystal
        t0 = time.perf_counter()
        self._check(fn(self._h, None, 0, 0))
        return (time.perf_counter() - t0) * 1e3

    def load_bitmap_crystal(self, words: np.ndarray):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        self._check(self._L.smg_load_bitmap_crystal(
            self._h, words.ctypes.data_as(C.POINTER(C.c_uint32)), words.size))

    def load_bitmap_collimator(self, words: np.ndarray):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        self._check(self._L.smg_load_bitmap_collimator(
            self._h, words.ctypes.data_as(C.POINTER(C.c_uint32)), words.size))

    # ---- scatter -----------------------------------------------------------------
    def scatter_crystal(self, pe: np.ndarray, rotation: int = 0, precision: str = "fp32"):
        pe = np.ascontiguousarray(pe, dtype=np.float32).reshape(self.dims.n_det, self.dims.n_vox)
        out = np.empty((self.dims.n_det, self.dims.n_vox), np.float32)
        st = Stats()
        self._check(self._L.smg_scatter_crystal(self._h, rotation, PREC[precision],
                                                pe.ctypes.data_as(C.c_void_p), 0,
                                                out.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
        return out, st.as_dict()

    def scatter_collimator(self, rotation: int = 0, variant: str = "combined",
                           precision: str = "fp32"):
        out = np.empty((self.dims.n_det, self.dims.n_vox), np.float32)
        st = Stats()
        self._check(self._L.smg_scatter_collimator(self._h, rotation, COLL_VARIANT[variant],
                                                   PREC[precision],
                                                   out.ctypes.data_as(C.c_void_p), 0,
                                                   C.byref(st)))
        return out, st.as_dict()

    def scatter_sum(self, parts, pe_drive=None, pe_add=None, rotation: int = 0,
                    precision: str = "fp32", scatter: bool = True, total: bool = True):
        """Scatter block of `parts` ("crystal", "combined", "standalone"; the
        crystal part driven by pe_drive) and pe_add + scatter, both formed on
        the GPU.  Returns (scatter or None, sum or None, stats)."""
        bits = {"crystal": 1, "combined": 2, "standalone": 4}
        mask = 0
        for name in parts:
            mask |= bits[name]
        shape = (self.dims.n_det, self.dims.n_vox)

        def host(a):
            return None if a is None else np.ascontiguousarray(a, np.float32).reshape(shape)

        pe_drive, pe_add = host(pe_drive), host(pe_add)
        sc = np.empty(shape, np.float32) if scatter else None
        sm = np.empty(shape, np.float32) if total else None
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        st = Stats()
        self._check(self._L.smg_scatter_sum(self._h, rotation, mask, PREC[precision], ptr(pe_drive),
                                            ptr(pe_add), 0, ptr(sc), ptr(sm), 0, C.byref(st)))
        return sc, sm, st.as_dict()

    def scatter_collimator_grid(self, pitch_mm: float, rotation: int = 0,
                                precision: str = "fp32", layers: int = 1):
        """Collimator scatter from grid points over the tungsten of the layers
        in the bit mask `layers` (1: the layer-0 plate).
        Returns (matrix [N][V], stats, number of grid points)."""
        out = np.empty((self.dims.n_det, self.dims.n_vox), np.float32)
        st = Stats()
        npts = C.c_int64(0)
        self._check(self._L.smg_scatter_collimator_grid(self._h, rotation, float(pitch_mm),
                                                        int(layers), PREC[precision],
                                                        out.ctypes.data_as(C.c_void_p), 0,
                                                        C.byref(npts), C.byref(st)))
        return out, st.as_dict(), npts.value

    def scatter_crystal_grid(self, grid, rows=None, rotation: int = 0, precision: str = "fp32"):
        """Crystal scatter from the centres of a grid (gx, gy, gz) of cells
        in every bar, for detector rows `rows` (all bars when None).
        Returns (matrix [len(rows)][V], stats)."""
        if rows is None:
            n = self.dims.n_det
            rp = None
        else:
            r = np.ascontiguousarray(rows, np.int32)
            n = len(r)
            rp = r.ctypes.data_as(C.c_void_p)
        out = np.empty((n, self.dims.n_vox), np.float32)
        st = Stats()
        self._check(self._L.smg_scatter_crystal_grid(
            self._h, rotation, *grid, rp, n if rows is not None else 0, PREC[precision],
            out.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
        return out, st.as_dict()

def open_context(p, device: int = 0) -> Context:
    """Context with the geometry of ``p`` (a :class:`params.Params`)."""
    ctx = Context(device)
    ctx.set_geometry(p)
    return ctx
