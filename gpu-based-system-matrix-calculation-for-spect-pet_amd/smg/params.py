"""Parameter-file contract of the SPECT/PET system-matrix generator.

The four geometry inputs are raw little-endian float32 arrays read from the
working directory under fixed names and fixed lengths
(PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:33-56):

    Params_Collimator.dat  80000 floats
    Params_Detector.dat    80000 floats
    Params_Image.dat         100 floats
    Params_Physics.dat       100 floats

Field map (PEGen_RayTracing_CircularHole/ReadMe.txt:1-61, and the reads in
PEGen_RayTracing_CircularHole/PESysMatGen.cu:958-1075, 1240-1271):

    coll[0]                     number of collimator layers Lc
    coll[(m+1)*10 + 0..7]       holes in layer m, width (x), thickness (y),
                                height (z), distance from layer 0 (y),
                                mu_total, mu_pe, mu_compton  (mu in 1/mm)
    coll[100 + 9*h + 0..8]      hole h (global index over layers): x, y1, y2,
                                z, radius, mu_total, mu_pe, mu_compton, flag
    det[0]                      number of detector bars N
    det[12*d + 1..12]           bar d: x, y, z, width (x), thickness (y),
                                height (z), mu_total, mu_pe, mu_compton,
                                energy resolution (FWHM fraction), rotation
                                about y (rad), flag
    img[0..11]                  NX, NY, NZ, voxel widths x/y/z, rotations,
                                angle per rotation, FOV shift x/y/z,
                                FOV-centre-to-collimator distance F
    img[20]                     rotation index (written per rotation)
    phys[0..9]                  Compton on, save PE, save scatter, save sum,
                                same energy window, window low/high (keV),
                                photopeak energy E0 (keV), compute crystal
                                bitmap, compute collimator bitmap

All lengths are in mm and every attenuation coefficient is per mm: the
reference multiplies mu by mm path lengths (PESysMatGen.cu:1278-1283).

The seeded generators below build the five benchmark geometries.  Choices the
benchmark description leaves open are fixed here once and listed in
``GENERATOR_NOTES``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

N_COLL = 80000
N_DET = 80000
N_IMG = 100
N_PHYS = 100

FILE_NAMES = {
    "coll": "Params_Collimator.dat",
    "det": "Params_Detector.dat",
    "img": "Params_Image.dat",
    "phys": "Params_Physics.dat",
}
FILE_LENGTHS = {"coll": N_COLL, "det": N_DET, "img": N_IMG, "phys": N_PHYS}

# Capacities implied by the fixed array lengths.
MAX_LAYERS = 9            # (m+1)*10+7 must stay below the hole table at 100
MAX_HOLES = (N_COLL - 100) // 9
MAX_BARS = (N_DET - 1) // 12

GENERATOR_NOTES = {
    "F": "FOV-centre-to-collimator distance F = 30 mm for every config",
    "mu_units": "mu in 1/mm: 36.5/cm tungsten -> 3.65/mm, 2.5/cm bar -> 0.25/mm",
    "coll_mu": "tungsten mu_pe = 3.30/mm, mu_compton = 0.35/mm (mu_total 3.65)",
    "bar_mu": "bar mu_total 0.25, mu_pe 0.18, mu_compton 0.07 /mm",
    "resolution": "bar energy resolution 0.10 FWHM at E0",
    "cfg1_holes": "x, z = default_rng(0).uniform(-29.5, 29.5, 16) (x drawn first)",
    "cfg2_holes": "20x20 grid pitch 4 mm (+-38), default_rng(1): jitter x, jitter z"
                  " ~ U(-0.2, 0.2), then radius ~ U(0.4, 0.6); x index fastest",
    "cfg1_bars": "8x8 bars 2x10x2 mm at x,z in {-7..7 step 2}, y = 17.5; x fastest",
    "cfg2_bars": "2 layers of 32x32 bars 2.5x8x2.5 mm, pitch 2.5, y = 18 and 26",
    "window": "same window for all bars: 126.45-154.55 keV, E0 = 140.5 keV",
    "cfg3_image": "cfg1 geometry with a 32x32x16 image at 1 mm",
    "cfg4_image": "cfg2 detector and collimator with a 32x32x1 image at 1 mm",
}


@dataclass
class Params:
    """The four float32 parameter arrays plus generator metadata."""

    coll: np.ndarray
    det: np.ndarray
    img: np.ndarray
    phys: np.ndarray
    meta: dict = field(default_factory=dict)

    @classmethod
    def empty(cls) -> "Params":
        return cls(np.zeros(N_COLL, np.float32), np.zeros(N_DET, np.float32),
                   np.zeros(N_IMG, np.float32), np.zeros(N_PHYS, np.float32))

    # ---- derived sizes, parsed the way the reference parses them -------------
    @property
    def n_det(self) -> int:
        return int(np.floor(self.det[0] + np.float32(0.0001)))

    @property
    def n_layers(self) -> int:
        return int(np.floor(self.coll[0] + np.float32(0.001)))

    @property
    def dims(self) -> tuple[int, int, int]:
        return tuple(int(np.floor(self.img[i] + np.float32(0.001))) for i in range(3))

    @property
    def n_vox(self) -> int:
        nx, ny, nz = self.dims
        return nx * ny * nz

    @property
    def n_rot(self) -> int:
        return int(np.floor(self.img[6] + 0.001))

    @property
    def n_holes_layer0(self) -> int:
        return int(np.floor(self.coll[10] + np.float32(0.01)))

    def holes_in_layer(self, m: int) -> int:
        return int(np.floor(self.coll[(m + 1) * 10]))

    @property
    def n_holes(self) -> int:
        return sum(self.holes_in_layer(m) for m in range(self.n_layers))

    def copy(self) -> "Params":
        return Params(self.coll.copy(), self.det.copy(), self.img.copy(),
                      self.phys.copy(), dict(self.meta))

    # ---- file I/O ------------------------------------------------------------
    def write(self, directory: str) -> None:
        os.makedirs(directory, exist_ok=True)
        for key, name in FILE_NAMES.items():
            arr = np.ascontiguousarray(getattr(self, key), dtype="<f4")
            assert arr.size == FILE_LENGTHS[key]
            arr.tofile(os.path.join(directory, name))

    @classmethod
    def read(cls, directory: str) -> "Params":
        arrays = {}
        for key, name in FILE_NAMES.items():
            path = os.path.join(directory, name)
            arr = np.fromfile(path, dtype="<f4")
            if arr.size < FILE_LENGTHS[key]:
                raise ValueError(f"{path}: {arr.size} floats, expected {FILE_LENGTHS[key]}")
            arrays[key] = arr[: FILE_LENGTHS[key]].astype(np.float32)
        return cls(**arrays)


# ---- field setters ---------------------------------------------------------------

def set_layer(p: Params, m: int, n_holes: int, width: float, thickness: float,
              height: float, dist: float, mu_t: float, mu_pe: float, mu_c: float) -> None:
    base = (m + 1) * 10
    p.coll[base:base + 8] = [n_holes, width, thickness, height, dist, mu_t, mu_pe, mu_c]


def set_hole(p: Params, h: int, x: float, y1: float, y2: float, z: float, r: float,
             mu_t: float = 0.0, mu_pe: float = 0.0, mu_c: float = 0.0, flag: float = 1.0) -> None:
    base = 100 + 9 * h
    p.coll[base:base + 9] = [x, y1, y2, z, r, mu_t, mu_pe, mu_c, flag]


def set_bar(p: Params, d: int, x: float, y: float, z: float, w: float, t: float, h: float,
            mu_t: float, mu_pe: float, mu_c: float, res: float, beta: float = 0.0,
            flag: float = 1.0) -> None:
    base = 12 * d
    p.det[base + 1:base + 13] = [x, y, z, w, t, h, mu_t, mu_pe, mu_c, res, beta, flag]


def set_image(p: Params, nx: int, ny: int, nz: int, wx: float, wy: float, wz: float,
              n_rot: int = 1, dphi: float = 0.0, shift=(0.0, 0.0, 0.0), fov2coll: float = 30.0) -> None:
    p.img[0:12] = [nx, ny, nz, wx, wy, wz, n_rot, dphi, shift[0], shift[1], shift[2], fov2coll]
    p.img[20] = 0.0


def set_physics(p: Params, compton: int = 1, save_pe: int = 1, save_scatter: int = 1,
                save_sum: int = 1, same_window: int = 1, lo: float = 126.45,
                hi: float = 154.55, e0: float = 140.5, calc_crystal_bitmap: int = 1,
                calc_coll_bitmap: int = 1) -> None:
    p.phys[0:10] = [compton, save_pe, save_scatter, save_sum, same_window, lo, hi, e0,
                    calc_crystal_bitmap, calc_coll_bitmap]


# ---- benchmark geometries ----------------------------------------------------------

TUNGSTEN = dict(mu_t=3.65, mu_pe=3.30, mu_c=0.35)
BAR_MU = dict(mu_t=0.25, mu_pe=0.18, mu_c=0.07)


def _cfg1_base() -> Params:
    p = Params.empty()
    p.coll[0] = 1
    set_layer(p, 0, 16, 60.0, 5.0, 60.0, 0.0, **TUNGSTEN)
    rng = np.random.default_rng(0)
    xs = rng.uniform(-29.5, 29.5, 16)
    zs = rng.uniform(-29.5, 29.5, 16)
    for h in range(16):
        set_hole(p, h, xs[h], -2.5, 2.5, zs[h], 0.5)
    p.det[0] = 64
    for iz in range(8):
        for ix in range(8):
            set_bar(p, iz * 8 + ix, -7.0 + 2.0 * ix, 17.5, -7.0 + 2.0 * iz, 2.0, 10.0, 2.0,
                    res=0.10, **BAR_MU)
    set_image(p, 32, 32, 1, 1.0, 1.0, 1.0)
    set_physics(p)
    return p


def _cfg2_base() -> Params:
    p = Params.empty()
    p.coll[0] = 1
    set_layer(p, 0, 400, 80.0, 8.0, 80.0, 0.0, **TUNGSTEN)
    rng = np.random.default_rng(1)
    jx = rng.uniform(-0.2, 0.2, 400)
    jz = rng.uniform(-0.2, 0.2, 400)
    rr = rng.uniform(0.4, 0.6, 400)
    for iz in range(20):
        for ix in range(20):
            h = iz * 20 + ix
            set_hole(p, h, -38.0 + 4.0 * ix + jx[h], -4.0, 4.0, -38.0 + 4.0 * iz + jz[h], rr[h])
    p.det[0] = 2048
    for layer, y in enumerate((18.0, 26.0)):
        for iz in range(32):
            for ix in range(32):
                set_bar(p, layer * 1024 + iz * 32 + ix, (ix - 15.5) * 2.5, y, (iz - 15.5) * 2.5,
                        2.5, 8.0, 2.5, res=0.10, **BAR_MU)
    set_image(p, 64, 64, 32, 0.5, 0.5, 0.5)
    set_physics(p)
    return p


def config(idx: int) -> Params:
    """Geometry of benchmark config ``idx`` (1..5), with its sampling in ``meta``."""
    if idx == 1:
        p = _cfg1_base()
        p.meta = dict(name="cfg1_pe_32x32x1_64bars", subsamples=(4, 4, 4))
    elif idx == 2:
        p = _cfg2_base()
        p.meta = dict(name="cfg2_pe_64x64x32_2048bars", subsamples=(3, 3, 3))
    elif idx == 3:
        p = _cfg1_base()
        set_image(p, 32, 32, 16, 1.0, 1.0, 1.0)
        p.meta = dict(name="cfg3_collscatter_32x32x16_64bars", subsamples=(1, 1, 1),
                      pe_subsamples=(4, 4, 4))
    elif idx == 4:
        p = _cfg2_base()
        set_image(p, 32, 32, 1, 1.0, 1.0, 1.0)
        p.meta = dict(name="cfg4_crystalscatter_32x32x1_2048bars", subsamples=(1, 1, 1),
                      pe_subsamples=(3, 3, 3))
    elif idx == 5:
        p = _cfg1_base()
        p.meta = dict(name="cfg5_adjoint_32x32x1_64bars", subsamples=(4, 4, 4))
    else:
        raise ValueError(f"unknown config {idx}")
    return p


def large_detector(n_layers: int = 3, nx: int = 50, nz: int = 40) -> Params:
    """The full-size collimator with a larger detector for the bitmap
    workloads: n_layers layers of nx x nz bars 2.5x8x2.5 mm at pitch 2.5 mm,
    y = 18, 26, 34, ... (3 x 50 x 40 = 6,000 bars by default, the size of the
    crystal bitmap's N^3 pass the reference describes: 13.5 GB of bits)."""
    p = _cfg2_base()
    n = n_layers * nx * nz
    if n > MAX_BARS:
        raise ValueError(f"{n} bars exceed the parameter layout's {MAX_BARS}")
    p.det[:] = 0.0
    p.det[0] = n
    for layer in range(n_layers):
        for iz in range(nz):
            for ix in range(nx):
                set_bar(p, (layer * nz + iz) * nx + ix, (ix - (nx - 1) / 2) * 2.5, 18.0 + 8.0 * layer,
                        (iz - (nz - 1) / 2) * 2.5, 2.5, 8.0, 2.5, res=0.10, **BAR_MU)
    p.meta = dict(name=f"large_detector_{n}bars")
    return p


def large_fov_pe() -> Params:
    """The system size the reference's README quotes a PE time for (ReadMe.md:
    200): a 200 x 200 2D field of view (0.5 mm voxels, one slice) and 6,000
    bars (:func:`large_detector`), at the reference's default 8 x 16 x 8
    sub-cells (PESysMatGen.cu:1037): 240,000,000 elements, 2.46e11 sub-cell
    rays."""
    p = large_detector()
    set_image(p, 200, 200, 1, 0.5, 0.5, 0.5)
    p.meta = dict(name="pe_200x200x1_6000bars", subsamples=(8, 16, 8))
    return p


def synthetic_mu_tables(power_law: bool = True, e0: float = 140.5):
    """mu(E) tables of the scatter workloads: 64 log-spaced nodes on 20-200
    keV, returned as (energy_kev, crystal_pe, crystal_compton, coll_pe,
    coll_compton) in 1/mm.  power_law: the configs' mu at e0 scaled as
    (e0/E)^3 (photoelectric) and e0/E (Compton), which log-log interpolation
    reproduces exactly; otherwise smooth tables with other slopes and a
    modulation."""
    E = 20.0 * 10.0 ** (np.arange(64) / 63.0)
    r = e0 / E
    cpe, cc = BAR_MU["mu_pe"], BAR_MU["mu_c"]
    wpe, wc = TUNGSTEN["mu_pe"], TUNGSTEN["mu_c"]
    if power_law:
        return E, cpe * r ** 3, cc * r, wpe * r ** 3, wc * r
    lE = np.log(E)
    return (E, cpe * r ** 2.7 * (1 + 0.08 * np.sin(3 * lE)), cc * r ** 0.85 * (1 + 0.05 * np.cos(2 * lE)),
            wpe * r ** 2.9 * (1 + 0.06 * np.sin(2.5 * lE)), wc * r ** 1.1)


def grad_param_count(p: Params) -> int:
    """Gradient vector length: one mu per collimator layer, then mu_total and
    mu_pe of every bar."""
    return p.n_layers + 2 * p.n_det
