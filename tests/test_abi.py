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


def test_vtu_writer_round_trip(tmp_path):
    """The VTU output parses as XML with the reference's layout: padded
    coordinates, connectivity in element order, offsets, VTK types, 2D point
    vectors padded to 3 components, cell data."""
    import xml.etree.ElementTree as ET
    import numpy as np
    from statfem_amd import io, mesh as M
    m = M.grid2d(3, 2)
    u = np.arange(m.nnodes, dtype=float)
    g = np.stack([u, -u], axis=1)
    k = np.linspace(1, 2, m.nel)
    path = tmp_path / "out.vtu"
    io.write_vtu(str(path), m, point_data={"u": u, "grad": g}, cell_data={"kappa": k})
    root = ET.parse(str(path)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == m.nnodes and int(piece.get("NumberOfCells")) == m.nel

    def arr(node, name):
        for d in node.iter("DataArray"):
            if d.get("Name") == name:
                return d, np.array(d.text.split(), float)
        raise KeyError(name)

    d, xyz = arr(piece, "Coordinates")
    assert d.get("NumberOfComponents") == "3"
    assert np.allclose(xyz.reshape(-1, 3)[:, :2], m.coords) and np.all(xyz.reshape(-1, 3)[:, 2] == 0)
    _, c = arr(piece, "connectivity")
    assert np.array_equal(c.reshape(m.nel, 4), m.conn)
    _, off = arr(piece, "offsets")
    assert np.array_equal(off, 4 * np.arange(1, m.nel + 1))
    _, ty = arr(piece, "types")
    assert np.all(ty == 9)
    d, gg = arr(piece.find("PointData"), "grad")
    assert d.get("NumberOfComponents") == "3" and np.allclose(gg.reshape(-1, 3)[:, :2], g)
    _, kk = arr(piece.find("CellData"), "kappa")
    assert np.allclose(kk, k)


def test_launch_shape_cache_round_trip(tmp_path, monkeypatch):
    """The autotuner's persisted choices: keys separate devices, shapes,
    transposition and batch layouts; a stored choice is read back whole, and
    a second store keeps the first."""
    import numpy as np
    from statfem_amd import autotune
    path = str(tmp_path / "sub" / "shapes.json")
    monkeypatch.setenv("STATFEM_AMD_TUNE_CACHE", path)
    assert autotune.cache_path() == path
    assert autotune.load_cache() == {}
    info = {"arch": "gfx950:sramecc+:xnack-", "num_cu": 256}
    seg = np.array([0, 100, 250], np.int32)
    k1 = autotune.problem_key(info, 250, 127, True, seg, 1e-10)
    keys = {k1, autotune.problem_key(info, 250, 127, False, seg, 1e-10),
            autotune.problem_key(info, 250, 126, True, seg, 1e-10),
            autotune.problem_key(info, 250, 127, True, np.array([0, 101, 250]), 1e-10),
            autotune.problem_key(dict(info, num_cu=304), 250, 127, True, seg, 1e-10),
            autotune.problem_key(info, 250, 127, True, seg, 1e-12)}
    assert len(keys) == 6
    assert k1 == autotune.problem_key(info, 250, 127, True, seg.astype(np.int64), 1e-10)
    autotune.store_choice(k1, {"lanes": 8, "rhs_tile": 0, "ms": 1.5})
    k2 = autotune.problem_key(info, 90, 4, False, None, 1e-10)
    autotune.store_choice(k2, {"lanes": 1, "rhs_tile": 0, "ms": 0.2})
    d = autotune.load_cache()
    assert d[k1] == {"lanes": 8, "rhs_tile": 0, "ms": 1.5} and d[k2]["lanes"] == 1
    with open(path, "w") as f:   # a damaged file reads as empty, and is replaced by the next store
        f.write("{not json")
    assert autotune.load_cache() == {}
    autotune.store_choice(k1, {"lanes": 16, "rhs_tile": 64})
    assert autotune.load_cache() == {k1: {"lanes": 16, "rhs_tile": 64}}
