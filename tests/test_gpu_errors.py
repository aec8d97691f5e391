"""Error behaviour of the C ABI: invalid arguments and call order return a
status and a message (the reference exits or proceeds unchecked,
PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:39-41), and the context
stays usable after an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BAD_PARAM, CAPACITY, STATE = 1, 2, 6


def status_of(exc_info):
    return exc_info.value.status


def test_call_order_and_arguments(smg_mod, params_mod):
    L = smg_mod.lib()
    ctx = smg_mod.Context(0)
    try:
        out = np.zeros(4, np.float32)
        # no geometry yet
        s = L.smg_pe_forward(ctx._h, 0, 0, out.ctypes.data_as(C.c_void_p), 0, None)
        assert s == STATE and b"no geometry" in L.smg_last_error(ctx._h)
        p = params_mod.config(1)
        ctx.set_geometry(p)
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.set_pe_subsamples(0, 1, 1)
        assert status_of(e) == BAD_PARAM
        # negative rotation, NULL output
        assert L.smg_pe_forward(ctx._h, -1, 0, out.ctypes.data_as(C.c_void_p), 0, None) == BAD_PARAM
        assert L.smg_pe_forward(ctx._h, 0, 0, None, 0, None) == BAD_PARAM
        assert b"NULL" in L.smg_last_error(ctx._h)
        assert L.smg_pe_project(ctx._h, 0, 0, None, 0, out.ctypes.data_as(C.c_void_p), 0,
                                None) == BAD_PARAM
        # grid scatter arguments
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.scatter_collimator_grid(0.0)
        assert status_of(e) == BAD_PARAM
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.scatter_collimator_grid(1.0, layers=0)
        assert status_of(e) == BAD_PARAM
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.scatter_crystal_grid((2, 2, 2), rows=[0, 64])
        assert status_of(e) == BAD_PARAM and "out of range" in str(e.value)
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.scatter_crystal_grid((100, 100, 10))
        assert status_of(e) == BAD_PARAM
        # mu(E) tables: energies must increase
        e_kev = np.array([20.0, 10.0, 200.0])
        v = np.ones(3)
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.set_mu_tables(e_kev, v, v, v, v)
        assert status_of(e) == BAD_PARAM and "increasing" in str(e.value)
        # a cached bitmap that is too short
        with pytest.raises(smg_mod.SmgError) as e:
            ctx.load_bitmap_crystal(np.zeros(10, np.uint32))
        assert status_of(e) == BAD_PARAM
        # still usable: the same matrix as a fresh context
        ctx.set_pe_subsamples(2, 2, 2)
        a, _ = ctx.pe_forward(0, "fp32")
    finally:
        ctx.close()
    with smg_mod.open_context(p) as fresh:
        fresh.set_pe_subsamples(2, 2, 2)
        b, _ = fresh.pe_forward(0, "fp32")
    assert np.array_equal(a, b)


def test_bad_device_index(smg_mod):
    h = C.c_void_p()
    s = smg_mod.lib().smg_create(1000, C.byref(h))
    assert s != 0 and not h.value


def test_reference_shim_refuses_fractional_rotation(params_mod, tmp_path):
    """The reference multiplies img[20], as a float, by the angle step
    (PEGen_RayTracing_CircularHole/PESysMatGen.cu:1004-1006); the C ABI takes
    whole rotation indices, so the PESysMatGen shim refuses any other value
    with a message and a failing exit status instead of rounding it."""
    import os
    import subprocess
    import sys
    import textwrap

    from conftest import PKG
    p = params_mod.config(1)
    p.img[20] = 0.5
    p.write(str(tmp_path))
    code = textwrap.dedent(f"""
        import ctypes as C, numpy as np, sys
        sys.path.insert(0, {PKG!r})
        import smg
        L = smg.lib()
        f = getattr(L, "_Z11PESysMatGenPfS_S_S_i")
        arrs = [np.fromfile({str(tmp_path)!r} + "/Params_" + n + ".dat", np.float32)
                for n in ("Collimator", "Detector", "Image")]
        dst = np.zeros(64 * 1024, np.float32)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        f(ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), ptr(dst), 0)
        print("returned")
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "rotation index" in r.stderr
