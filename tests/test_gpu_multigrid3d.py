"""Multigrid-preconditioned solves on declared 3D grids (hex8, 27-point block
stencils) against the CPU FP64 oracle's direct solve: vector elasticity with a
clamped face and scalar operators, odd and even node counts, ragged and
one-column right-hand sides, both orientations."""
import numpy as np
import pytest

from statfem_amd import mesh as M

pytestmark = pytest.mark.gpu


def _elastic(orc, nx, ny, nz, seed):
    m = M.grid3d(nx, ny, nz)
    E = np.exp(0.4 * np.random.default_rng(seed).normal(size=m.nel))
    bcn = np.where(m.coords[:, 2] < 1e-9)[0]
    bc = np.sort(np.concatenate([3 * bcn + c for c in range(3)])).astype(np.int32)
    rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, E, 0.3, nq1d=2, bc_dofs=bc)
    return m, E, bc, rp, ci, va


def _elastic_system(ctx, m, E, bc):
    from statfem_amd import hip
    s = hip.SystemSolveHip(ctx, m, 3)
    s.set_grid()
    s.assemble(hip.ELASTICITY, E, 0.3, nq1d=2)
    s.applyConstraints(bc)
    return s


@pytest.mark.parametrize("trans", [False, True])
def test_elasticity_multigrid_matches_direct(ctx, orc, trans):
    from statfem_amd import hip
    m, E, bc, rp, ci, va = _elastic(orc, 7, 6, 9, seed=1)   # 8 x 7 x 10 nodes, 1680 DOFs
    rng = np.random.default_rng(2)
    k = 21
    B = rng.normal(size=(3 * m.nnodes, k))
    B[:, 4] = 0.0
    s = _elastic_system(ctx, m, E, bc)
    X, st = s.solve(B, trans=trans, opts=hip.default_opts(tol=1e-12))
    Xr = orc.LU(rp, ci, va).solve(B, trans=trans)
    Xg = X.download()
    err = np.linalg.norm(Xg - Xr, axis=0) / np.maximum(np.linalg.norm(Xr, axis=0), 1e-300)
    assert st["layout"] == 4 and st["precond"] == 2 and st["levels"] >= 2 and st["converged"] == 1, st
    assert err.max() < 1e-8, (err.max(), st)
    assert np.all(Xg[:, 4] == 0.0)
    assert st["inner_iters"] < 80, st


def test_elasticity_one_column_and_jacobi_fallback(ctx, orc):
    from statfem_amd import hip
    m, E, bc, rp, ci, va = _elastic(orc, 8, 8, 8, seed=3)
    s = _elastic_system(ctx, m, E, bc)
    b = np.random.default_rng(4).normal(size=(3 * m.nnodes, 1))
    ref = orc.LU(rp, ci, va).solve(b)
    X, st = s.solve(b, opts=hip.default_opts(tol=1e-11))
    assert st["layout"] == 4 and st["precond"] == 2, st
    assert np.abs(X.download() - ref).max() <= 1e-8 * np.abs(ref).max()
    X, st = s.solve(b, opts=hip.default_opts(tol=1e-11, precond=1))
    assert st["layout"] != 4 and st["precond"] == 1, st
    assert np.abs(X.download() - ref).max() <= 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("nu,max_iters", [(0.3, 25), (0.45, 40)])
def test_elasticity_block_smoother_weights(ctx, orc, nu, max_iters):
    """Nodal block-Jacobi smoothing weighted per level by a power-iteration
    estimate of lambda_max(B^-1 A): 15^3 cells (16 nodes per axis, so the first
    coarse level has half-width last cells, where lambda_max reaches 3.6 at
    nu = 0.3) and nu = 0.45 (lambda_max near 4), where a fixed weight of 0.6
    would diverge. Point l1-Jacobi took 28-29 and 49-52 iterations here in a
    CPU model of the cycle; the bounds pin the block smoother's gain."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from statfem_amd import hip
    m = M.grid3d(15, 15, 15)
    E = np.exp(0.4 * np.random.default_rng(5).normal(size=m.nel))
    bcn = np.where(m.coords[:, 2] < 1e-9)[0]
    bc = np.sort(np.concatenate([3 * bcn + c for c in range(3)])).astype(np.int32)
    rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, E, nu, nq1d=2, bc_dofs=bc)
    s = hip.SystemSolveHip(ctx, m, 3)
    s.set_grid()
    s.assemble(hip.ELASTICITY, E, nu, nq1d=2)
    s.applyConstraints(bc)
    B = np.random.default_rng(6).normal(size=(3 * m.nnodes, 8))
    B[bc] = 0.0
    X, st = s.solve(B, opts=hip.default_opts(tol=1e-10))
    A = sp.csc_matrix(sp.csr_matrix((va, ci, rp), shape=(3 * m.nnodes,) * 2))
    Xr = spl.splu(A).solve(B)
    err = np.linalg.norm(X.download() - Xr, axis=0) / np.linalg.norm(Xr, axis=0)
    assert st["layout"] == 4 and st["precond"] == 2 and st["converged"] == 1, st
    assert err.max() < 1e-7, (err.max(), st)   # residual 1e-10 at a condition number near 1e4
    assert st["inner_iters"] <= max_iters, st


@pytest.mark.parametrize("physics", ["poisson", "matern"])
def test_scalar_3d_multigrid_matches_direct(ctx, orc, physics):
    from statfem_amd import hip
    m = M.grid3d(10, 9, 8)
    X = m.coords
    bc = np.where((X[:, 0] < 1e-9) | (X[:, 1] > 1 - 1e-9))[0].astype(np.int32)
    kap = np.exp(0.5 * np.random.default_rng(5).normal(size=m.nel))
    s = hip.SystemSolveHip(ctx, m)
    s.set_grid()
    if physics == "poisson":
        s.assemble(hip.POISSON, kap, nq1d=2)
        rp, ci, va, _ = orc.assemble(m, orc.POISSON, kap, nq1d=2, bc_dofs=bc)
    else:
        s.assemble(hip.MATERN, 3.0, 0.8, nq1d=2)
        rp, ci, va, _ = orc.assemble(m, orc.MATERN, 3.0, 0.8, nq1d=2, bc_dofs=bc)
    s.applyConstraints(bc)
    B = np.random.default_rng(6).normal(size=(m.nnodes, 33))
    for trans in (False, True):
        Xd, st = s.solve(B, trans=trans, opts=hip.default_opts(tol=1e-12))
        Xr = orc.LU(rp, ci, va).solve(B, trans=trans)
        err = np.linalg.norm(Xd.download() - Xr, axis=0) / np.linalg.norm(Xr, axis=0)
        assert st["layout"] == 4 and st["precond"] == 2 and st["converged"] == 1, st
        assert err.max() < 1e-8, (err.max(), st)


def test_elasticity_multigrid_on_anisotropic_cells(ctx, orc):
    """Cells 4 : 1 : 1 (a 24 x 6 x 6 hex grid of the unit cube): slower
    smoothing, same answer as the direct solve."""
    from statfem_amd import hip
    m, E, bc, rp, ci, va = _elastic(orc, 24, 6, 6, seed=7)
    s = _elastic_system(ctx, m, E, bc)
    B = np.random.default_rng(8).normal(size=(3 * m.nnodes, 8))
    X, st = s.solve(B, opts=hip.default_opts(tol=1e-11))
    Xr = orc.LU(rp, ci, va).solve(B)
    assert st["layout"] == 4 and st["converged"] == 1, st
    assert np.abs(X.download() - Xr).max() <= 1e-8 * np.abs(Xr).max(), st


def test_elasticity_residual_at_scale(ctx, orc):
    """A larger clamped block (32^3 cells, 107,811 DOFs) with 16 right-hand
    sides, both orientations: the FP64 residual of every column against the
    oracle's own assembly of the operator meets the tolerance."""
    import scipy.sparse as sp
    from statfem_amd import hip
    m, E, bc, rp, ci, va = _elastic(orc, 32, 32, 32, seed=5)
    n = 3 * m.nnodes
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    s = _elastic_system(ctx, m, E, bc)
    B = np.random.default_rng(6).normal(size=(n, 16))
    B[bc] = 0.0
    for trans in (False, True):
        X, st = s.solve(B, trans=trans, opts=hip.default_opts(tol=1e-10))
        assert st["layout"] == 4 and st["precond"] == 2 and st["converged"] == 1, st
        X = X.download()
        R = B - (A.T @ X if trans else A @ X)
        rel = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
        assert rel.max() < 1e-9, (trans, rel.max())
