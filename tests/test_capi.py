"""The C ABI library loads and exports every entry point its headers declare
(include/smg.h, include/smg_reference.h); no GPU work is started."""
import os
import re
import subprocess

from conftest import PKG, ROOT


def declared_c_functions():
    text = open(os.path.join(ROOT, "include", "smg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smg_[a-z0-9_]+)\s*\(", text)))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_every_declared_c_entry_point_is_exported():
    syms = exported_symbols(os.path.join(PKG, "libsmg.so"))
    names = declared_c_functions()
    assert len(names) >= 20
    missing = [n for n in names if n not in syms]
    assert not missing, missing


def test_reference_signature_shims_are_exported():
    """PESysMatGen and the three scatter() overloads with C++ linkage."""
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(PKG, "libsmg.so")],
                         capture_output=True, text=True, check=True).stdout
    assert "PESysMatGen(float*, float*, float*, float*, int)" in out
    assert ("scatter(float*, float*, float*, float*, float*, char const*, char const*, float*, int)"
            in out)
    assert "scatter(float*, float*, float*, float*, float*, char const*, float*, int)" in out
    assert "scatter(float*, float*, float*, float*, char const*, float*, int)" in out


def test_python_binding_loads(smg_mod):
    L = smg_mod.lib()
    for name in smg_mod.EXPORTS:
        assert hasattr(L, name)
    assert L.smg_bitmap_crystal_words(64) == (64 * 65 // 2 * 64 + 31) // 32
    assert L.smg_bitmap_collimator_words(400, 2048) == 400 * 2048 * 2048 // 32
    assert L.smg_status_string(2) == b"capacity exceeded"
