"""End-to-end runs of the four drivers on the GPU: same files as the reference
executables (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:141-147,
ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:197-230,
ScatterGen_Collimator/ScatterGen_Collimator.cpp:183-214,
ScatterGen_Crystal/ScatterGen_Crystal.cpp:181-204)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, TOL_FP32, assert_parity

pytestmark = pytest.mark.gpu
BIN = os.path.join(PKG, "bin")


def run(exe, args, cwd):
    r = subprocess.run([os.path.join(BIN, exe)] + args, cwd=cwd, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r


def shifted(name, p):
    return f"{name}_shift_{p.img[8]:f}_{p.img[9]:f}_{p.img[10]:f}.sysmat"


@pytest.fixture(scope="module")
def run_dir(tmp_path_factory, params_mod):
    d = tmp_path_factory.mktemp("drivers")
    p = params_mod.config(1)
    params_mod.set_image(p, 32, 32, 1, 1.0, 1.0, 1.0, n_rot=2, dphi=0.2, shift=(0.25, 0.0, 0.0))
    p.write(str(d))
    return d, p


def test_pegen_writes_all_rotations(run_dir, oracle_mod):
    d, p = run_dir
    run("PEGen_CircularHole", ["-cuda", "0", "-sub", "2", "2", "2"], d)
    A = np.fromfile(d / shifted("PE_SysMat", p), np.float32).reshape(2, 64, 1024)
    o = oracle_mod.Oracle(p)
    rows = np.array([0, 27, 63], np.int32)
    for r in range(2):
        assert_parity(A[r][rows], o.pe_rows(rows, rot=r, sub=(2, 2, 2)), TOL_FP32, f"driver PE rot {r}")


def test_scatter_circularhole_files_and_cache(run_dir, smg_mod, oracle_mod):
    d, p = run_dir
    run("ScatterGen_CircularHole", ["-cuda", "0"], d)
    pe = np.fromfile(d / shifted("PE_SysMat", p), np.float32).reshape(2, 64, 1024)
    sc = np.fromfile(d / shifted("Scatter_SysMat", p), np.float32).reshape(2, 64, 1024)
    full = np.fromfile(d / shifted("SysMat_withScatter", p), np.float32).reshape(2, 64, 1024)
    np.testing.assert_array_equal(full, pe + sc)
    bc = np.fromfile(d / "GeometryRelationShip_Crystal2Crystal", np.uint32)
    bl = np.fromfile(d / "GeometryRelationShip_Collimator2Crystal", np.uint32)
    o = oracle_mod.Oracle(p)
    assert np.array_equal(bc, o.bitmap_crystal())
    assert np.array_equal(bl, o.bitmap_collimator())
    # the rotation-0 PE block drives every rotation, as in the reference driver
    dd = np.array([3, 30, 60])
    vv = np.array([100, 500, 900])
    for r in range(2):
        ref = o.crystal_scatter_pairs(pe[0], bc, dd, vv, rot=r) + o.collimator_scatter_pairs(bl, dd, vv, rot=r)
        assert_parity(sc[r][dd, vv], ref, TOL_FP32, f"driver scatter rot {r}")
    # reading the cached bitmaps gives the same matrices
    os.rename(d / shifted("Scatter_SysMat", p), d / "first.sysmat")
    run("ScatterGen_CircularHole", ["-GeoCrystal", "GeometryRelationShip_Crystal2Crystal",
                                    "-GeoCollimator", "GeometryRelationShip_Collimator2Crystal"], d)
    again = np.fromfile(d / shifted("Scatter_SysMat", p), np.float32)
    assert np.array_equal(again, np.fromfile(d / "first.sysmat", np.float32))


def test_standalone_scatter_drivers(run_dir, smg_mod):
    d, p = run_dir
    run("ScatterGen_Collimator", [], d)
    run("ScatterGen_Crystal", [], d)
    pe = np.fromfile(d / shifted("PE_SysMat", p), np.float32).reshape(2, 64, 1024)
    coll = np.fromfile(d / shifted("Scatter_SysMat", p), np.float32).reshape(2, 64, 1024)
    cry = np.fromfile(d / shifted("Scatter_Crystal_SysMat", p), np.float32).reshape(2, 64, 1024)
    with smg_mod.open_context(p) as ctx:
        ref_coll, _ = ctx.scatter_collimator(1, "standalone", "fp32")
        ref_cry, _ = ctx.scatter_crystal(pe[0], 1, "fp32")
    assert np.array_equal(coll[1], ref_coll)
    assert np.array_equal(cry[1], ref_cry)
    total = np.fromfile(d / shifted("SysMat", p), np.float32).reshape(2, 64, 1024)
    np.testing.assert_array_equal(total, pe + cry)  # the crystal driver ran last


def test_scatter_driver_pe_per_rotation(run_dir, smg_mod):
    """-PErot drives each rotation with its own PE block (the reference passes
    rotation 0's for all); the files are written one rotation block at a time
    and no partial file is left behind."""
    d, p = run_dir
    run("ScatterGen_Crystal", ["-PErot"], d)
    pe = np.fromfile(d / shifted("PE_SysMat", p), np.float32).reshape(2, 64, 1024)
    cry = np.fromfile(d / shifted("Scatter_Crystal_SysMat", p), np.float32).reshape(2, 64, 1024)
    with smg_mod.open_context(p) as ctx:
        for r in range(2):
            ref, _ = ctx.scatter_crystal(pe[r], r, "fp32")
            assert np.array_equal(cry[r], ref)
    assert not np.array_equal(pe[0], pe[1])
    assert not list(d.glob("*.part"))
