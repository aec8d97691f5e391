"""Parameter-file contract and the drivers' command-line handling (no GPU):
PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:33-98 and the scatter
drivers' argument loops (ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:106-141)."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(PKG, "bin")


def test_generated_geometries_are_pinned(params_mod, tmp_path):
    ref = json.load(open(os.path.join(GOLDEN, "params_sha256.json")))
    for cfg in range(1, 6):
        p = params_mod.config(cfg)
        p.write(str(tmp_path / str(cfg)))
        for name, digest in ref[str(cfg)].items():
            got = hashlib.sha256(open(tmp_path / str(cfg) / name, "rb").read()).hexdigest()
            assert got == digest, (cfg, name)


def test_params_roundtrip_and_layout(params_mod, tmp_path):
    p = params_mod.config(2)
    p.write(str(tmp_path))
    q = params_mod.Params.read(str(tmp_path))
    for key in ("coll", "det", "img", "phys"):
        assert np.array_equal(getattr(p, key), getattr(q, key))
    assert os.path.getsize(tmp_path / "Params_Collimator.dat") == 80000 * 4
    assert os.path.getsize(tmp_path / "Params_Image.dat") == 100 * 4
    assert q.n_det == 2048 and q.n_vox == 64 * 64 * 32 and q.n_layers == 1
    assert q.holes_in_layer(0) == 400 and q.n_holes_layer0 == 400
    # field positions (ReadMe.txt): layer 0 record at 10..17, bar d at 12d+1..12
    assert q.coll[12] == 8.0 and q.coll[15] == np.float32(3.65)
    assert q.det[12 * 5 + 4] == 2.5 and q.det[12 * 5 + 5] == 8.0 and q.det[12 * 5 + 6] == 2.5
    assert q.img[11] == 30.0


def test_oracle_reproduces_golden_cfg1(oracle_mod, params_mod):
    ref = np.load(os.path.join(GOLDEN, "cfg1_pe_fp64.npy"))
    got = oracle_mod.Oracle(params_mod.config(1)).pe_matrix(sub=(4, 4, 4))
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


def run(exe, args, cwd):
    return subprocess.run([os.path.join(BIN, exe)] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=60)


DRIVERS = ["PEGen_CircularHole", "ScatterGen_CircularHole", "ScatterGen_Collimator",
           "ScatterGen_Crystal"]


@pytest.mark.parametrize("exe", DRIVERS)
def test_driver_help_and_errors(exe, tmp_path, params_mod):
    r = run(exe, ["-h"], tmp_path)
    assert r.returncode == 0 and "Usage" in r.stdout
    # missing parameter files: non-zero exit (the reference read them unchecked)
    r = run(exe, [], tmp_path)
    assert r.returncode == 1 and "Params_Collimator.dat" in r.stderr
    # a short parameter file
    params_mod.config(1).write(str(tmp_path))
    with open(tmp_path / "Params_Detector.dat", "r+b") as f:
        f.truncate(1000)
    r = run(exe, [], tmp_path)
    assert r.returncode == 1 and "Params_Detector.dat" in r.stderr
    # an unknown flag
    params_mod.config(1).write(str(tmp_path))
    r = run(exe, ["-bogus"], tmp_path)
    assert r.returncode == 1 and "Unknown parameter" in r.stderr


@pytest.mark.parametrize("exe", DRIVERS[1:])
def test_scatter_driver_missing_pe_file(exe, tmp_path, params_mod):
    params_mod.config(1).write(str(tmp_path))
    r = run(exe, ["-PE", "nope.sysmat"], tmp_path)
    assert r.returncode == 1 and "nope.sysmat" in r.stderr


@pytest.mark.parametrize("exe", DRIVERS[1:])
def test_scatter_driver_short_pe_file(exe, tmp_path, params_mod):
    """A PE matrix with fewer rotation blocks than the image asks for is an
    error before any output file is opened."""
    p = params_mod.config(1)
    params_mod.set_image(p, 32, 32, 1, 1.0, 1.0, 1.0, n_rot=2, dphi=0.2)
    p.write(str(tmp_path))
    np.zeros(64 * 1024 + 5, np.float32).tofile(tmp_path / "short.sysmat")  # one block and a bit
    r = run(exe, ["-PE", "short.sysmat"], tmp_path)
    assert r.returncode == 1 and "short.sysmat holds 65541 floats, 131072 expected" in r.stderr
    assert sorted(x.name for x in tmp_path.iterdir() if x.suffix in (".sysmat", ".part")) == \
        ["short.sysmat"]


def test_bench_refuses_more_gpus_and_unknown_workloads():
    """bench.py runs on one GPU and names its secondary workloads; both
    refusals happen before any GPU work."""
    bench = os.path.join(ROOT, "bench.py")
    r = subprocess.run([sys.executable, bench, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
    r = subprocess.run([sys.executable, bench, "--workloads", "crystal_scatter,bogus"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "bogus" in r.stderr
