"""CPU checks of the C-ABI boundary: the library loads and exports every
symbol that include/statfem_hip.h declares (no device calls)."""
import ctypes
import os
import re

from .conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "statfem_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sfh_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    from statfem_amd import hip
    L = ctypes.CDLL(hip.LIB_PATH)
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings_and_version():
    from statfem_amd import hip
    L = hip.lib()
    assert L.sfh_status_string(0) == b"ok"
    assert L.sfh_status_string(-2) == b"negative Jacobian"
    assert b"gfx950" in L.sfh_version()


def test_python_mirror_binds_every_symbol():
    from statfem_amd import hip
    assert set(_declared()) == set(hip._SIGS)


def test_cpp_dropin_header_compiles_and_client_is_built():
    """include/statfem/SystemSolveHip.hpp compiles standalone with the host
    compiler, and the C++ client used by the GPU test is built and linked."""
    import subprocess
    hdr = os.path.join(ROOT, "include", "statfem", "SystemSolveHip.hpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                        "-x", "c++", hdr], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    from statfem_amd import build
    assert os.path.exists(build.CPP_CLIENT)
