"""Per-element insertion path: GPU Triplet sort-merge against the oracle's
Triplet, and the C++ SystemSolve drop-in client."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_triplet_pattern_and_values_match_oracle(ctx, orc):
    """Shuffled element insertion order, duplicate keys and explicit zeros."""
    import ctypes as C
    from statfem_amd import hip, mesh as M
    m = M.grid2d(9, 7, kind=M.TRI3)
    Ke = orc.element_matrices(m, orc.POISSON, 1.3, nq1d=1)
    edofs = m.element_dofs()
    rng = np.random.default_rng(4)
    order = rng.permutation(m.nel)
    rows = np.repeat(edofs[order], 3, axis=1).ravel().astype(np.int32)
    cols = np.tile(edofs[order], (1, 3)).ravel().astype(np.int32)
    vals = Ke[order].reshape(-1).copy()
    vals[::17] = 0.0                                    # explicit zeros stay in the pattern
    h = C.c_void_p()
    hip._check(hip.lib().sfh_system_create_triplets(ctx.h, m.nnodes, len(rows), hip._p(rows), hip._p(cols),
                                                    hip._p(vals), C.byref(h)))
    n = C.c_int()
    nnz = C.c_longlong()
    hip._check(hip.lib().sfh_system_info(h, C.byref(n), C.byref(nnz)))
    rp = np.zeros(m.nnodes + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    hip._check(hip.lib().sfh_system_pattern(h, hip._p(rp), hip._p(ci)))
    va = np.zeros(nnz.value)
    hip._check(hip.lib().sfh_matrix_values(h, hip._p(va)))
    ts = orc.TripletSystem(m.nnodes)
    ts.insert_elements(edofs[order], vals.reshape(len(order), 3, 3))
    rpo, cio, vao = ts.csr()
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    assert np.abs(va - vao).max() <= 1e-14 * np.abs(vao).max()
    # a mesh without a diagonal entry on some row is refused
    bad_rows = np.array([0, 1], np.int32)
    bad_cols = np.array([1, 1], np.int32)
    h2 = C.c_void_p()
    rc = hip.lib().sfh_system_create_triplets(ctx.h, 2, 2, hip._p(bad_rows), hip._p(bad_cols),
                                              hip._p(np.ones(2)), C.byref(h2))
    assert rc == -1
    hip.lib().sfh_system_destroy(h)


def test_cpp_system_solve_dropin_client():
    from statfem_amd import build
    r = subprocess.run([build.CPP_CLIENT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
