"""First-order perturbation (ForwardPerturbation) on the GPU against the oracle."""
import numpy as np
import pytest

from statfem_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,declare", [(6, 5, 0), (13, 12, 1)])
def test_fpert_matches_oracle(ctx, orc, nx, ny, declare):
    """Mean and covariance at the sensors against the oracle's restatement
    (dense A^-1, the double sum over parameter pairs), with and without the
    forcing covariance; the derivative blocks of the product path equal the
    oracle's element matrices."""
    from statfem_amd import hip
    from statfem_amd.fpert import ForwardPerturbationHip, conductivity_derivatives
    m = M.grid2d(nx, ny)
    bc = M.dirichlet_all_boundary(m)
    kap = 1.0 + 0.3 * np.sin(np.arange(m.nel))
    s = hip.SystemSolveHip(ctx, m)
    if declare:
        s.set_grid()
    s.assemble(hip.POISSON, kap, nq1d=2)
    s.assemble_rhs(nq1d=2, fconst=1.0)
    s.applyConstraints(bc)
    dofs, blocks = conductivity_derivatives(m, bc)
    Ke = orc.element_matrices(m, orc.POISSON, np.ones(m.nel), nq1d=2)
    mask = np.zeros(m.nnodes, bool)
    mask[bc] = True
    oblk = np.array([np.where(mask[m.conn[e]][:, None], 0.0, Ke[e]) for e in range(m.nel)])
    assert np.abs(blocks - oblk).max() <= 1e-13 * np.abs(oblk).max()
    rng = np.random.default_rng(nx)
    xs = rng.uniform(0.1, 0.9, size=(7, 2))
    cols, vals = M.sensor_rows(m, xs)
    cen = m.coords[m.conn].mean(axis=1)
    fp = ForwardPerturbationHip(s, cols, vals, dofs, blocks, cen, reg_factor=1e-4)
    F = orc.body_force(m, nq1d=2, fconst=1.0)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, kap, nq1d=2, bc_dofs=bc, F=F)
    cv = orc.body_force(m, nq1d=2, fconst=1.0)
    cv[bc] = 0.0
    P = M.dense_projection(m.nnodes, cols, vals)
    for sf in (0.0, 0.4):
        prm = {"forcing.sigma": sf, "forcing.length": 0.3, "material.sigma": 0.2, "material.length": 0.25}
        mean, cov, st = fp.updateSolution(prm, hip.default_opts(tol=1e-12))
        mref, cref, _ = orc.fpert_update(rp, ci, va, F, m.coords, cv, sf, 0.3, P, dofs, oblk, cen, 0.2, 0.25,
                                         reg_factor=1e-4)
        assert np.abs(mean - mref).max() <= 1e-9 * np.abs(mref).max()
        assert np.abs(cov - cref).max() <= 1e-8 * np.abs(cref).max(), (sf, np.abs(cov - cref).max())
        assert np.allclose(fp.giveProjectedCovarU(), cov)
    fp.close()
