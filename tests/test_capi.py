"""The C-ABI library loads and exports every symbol declared in include/*.h
(no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fz_\w+)\s*\(", txt, flags=re.M))
    return names


def test_header_declares_entry_points():
    names = declared()
    assert {"fz_create", "fz_configure", "fz_forward_dev", "fz_forward3d", "fz_gain_hsra"} <= names


def test_library_exports_all_declared_symbols():
    from firtez_dz_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(gpu.exported_symbols()) <= declared()


def test_binding_fails_loudly_without_library(monkeypatch):
    from firtez_dz_amd import gpu
    monkeypatch.setattr(gpu, "LIB_PATH", "/nonexistent/libfirtez_gpu.so")
    monkeypatch.setattr(gpu, "_lib", None)
    with pytest.raises(gpu.FzError):
        gpu.library()


def test_fortran_module_and_caller_link():
    """The FIRTEZ_GPU module compiles with amdflang and a Fortran host
    (tests/fortran/fz_caller.f90) links against it and the library, binding
    the C symbols of FORWARD3D, INVERT3D, GET_PERTURBATION, HESSIAN."""
    import shutil
    import subprocess
    from firtez_dz_amd import build, gpu
    if not os.path.exists(build.FLANG) or not os.path.exists(gpu.LIB_PATH):
        pytest.skip("amdflang or the library not available")
    build.build_fortran()
    assert os.path.exists(build.CALLER)
    if shutil.which("nm"):
        und = subprocess.run(["nm", "-u", build.CALLER], capture_output=True, text=True).stdout
        for sym in ("fz_forward3d", "fz_invert3d", "fz_get_perturbation", "fz_hessian", "fz_create"):
            assert sym in und, sym
    if shutil.which("readelf"):
        dyn = subprocess.run(["readelf", "-d", build.CALLER], capture_output=True, text=True).stdout
        assert "libfirtez_gpu.so" in dyn
