"""3D elasticity with an SPDE Matern body force (the 64^3 hex8 workload at a
small size): GPU pipeline against a dense composition of oracle operators."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_spde_elasticity_prior_and_posterior_match_oracle(ctx, orc):
    from statfem_amd import configs, hip, mesh as M
    from statfem_amd.spde_elasticity import SPDEElasticity
    cfg = configs.c3_inputs(seed=3, n=4, ns=6)
    m = cfg["mesh"]
    tight = hip.default_opts(tol=1e-13, rhs_tile=0)
    pipe = SPDEElasticity(ctx, cfg, opts=tight, opts_matern=tight)
    mean, Cu = pipe.prior()
    y, g, u = pipe.synthesize()
    post = pipe.posterior(mean, Cu, y, field=True)

    # oracle: dense composition
    F = orc.body_force(m, nq1d=2, dofpn=3, fconst=[0.0, 0.0, cfg["fbar"]])
    rp, ci, va, F = orc.assemble(m, orc.ELASTICITY, cfg["E"], cfg["nu"], nq1d=2, bc_dofs=cfg["bc"], F=F)
    Ae = orc.LU(rp, ci, va)
    kap = cfg["kappa"]
    rpm, cim, vam, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    Am = np.zeros((m.nnodes, m.nnodes))
    for i in range(m.nnodes):
        Am[i, cim[rpm[i]:rpm[i + 1]]] = vam[rpm[i]:rpm[i + 1]]
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    tau = orc.matern_tau(cfg["sigma_g"], kap, cfg["nu_m"], 3)
    Ainv = np.linalg.inv(Am)
    Cg = Ainv @ np.diag(mass) @ Ainv / (tau * kap ** 2) ** 2
    w = mass.copy()
    w[cfg["bottom"]] = 0.0
    N = 3 * m.nnodes
    B = np.zeros((N, m.nnodes))
    B[3 * np.arange(m.nnodes) + 2, np.arange(m.nnodes)] = w
    P = np.zeros((len(cfg["cols"]), N))
    for s in range(P.shape[0]):
        np.add.at(P[s], cfg["cols"][s], cfg["vals"][s])
    Rt = Ae.solve(P.T.copy(), trans=True)
    Cu_ref = Rt.T @ B @ Cg @ B.T @ Rt
    ubar = Ae.solve(F)
    mean_ref = P @ ubar
    assert np.abs(mean - mean_ref).max() <= 1e-9 * np.abs(mean_ref).max()
    assert np.abs(Cu - Cu_ref).max() <= 1e-8 * np.abs(Cu_ref).max()

    # the synthetic reading is the forward model at the sampled load
    assert np.abs(Ae.solve(F + B @ g) - u).max() <= 1e-9 * np.abs(u).max()

    Ks = orc.mismatch_cov(cfg["xs"], 0.0, 1.0, cfg["sigma_e"])
    Y = np.atleast_2d(y)
    pm, pc = orc.post_proj(Cu_ref, mean_ref, Y, Ks, 1.0)
    assert np.abs(post["mean"] - pm).max() <= 1e-8 * np.abs(pm).max()
    assert np.abs(post["var"] - np.diag(pc)).max() <= 1e-8 * np.abs(np.diag(pc)).max()
    sign, ld = np.linalg.slogdet(Cu_ref + Ks)
    assert sign > 0 and abs(post["logdet"] - ld) <= 1e-9 * abs(ld)
    CuFull = Ae.solve(B @ Cg @ B.T @ Ae.solve(np.eye(N), trans=True), trans=False)
    fm, _ = orc.post_woodbury(CuFull, ubar, P, Ks, Y, 1.0, want_cov=False)
    assert np.abs(post["field_mean"] - fm).max() <= 1e-8 * np.abs(fm).max()


def test_elasticity_multi_rhs_solves_match_oracle(ctx, orc):
    """Vector-DOF (interleaved) operator: forward and transposed multi-RHS
    solves of the clamped row-zero system against the direct solver."""
    from statfem_amd import configs, hip
    cfg = configs.c3_inputs(seed=4, n=5, ns=3)
    m = cfg["mesh"]
    s = hip.SystemSolveHip(ctx, m, dofpn=3)
    s.assemble(hip.ELASTICITY, 2.0, 0.25, nq1d=2)
    s.applyConstraints(cfg["bc"])
    rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, 2.0, 0.25, nq1d=2, bc_dofs=cfg["bc"])
    lu = orc.LU(rp, ci, va)
    rng = np.random.default_rng(9)
    B = rng.normal(size=(s.ndof, 7))
    for trans in (False, True):
        X, st = s.solve(B, trans=trans, opts=hip.default_opts(tol=1e-12))
        ref = lu.solve(B, trans=trans)
        assert np.abs(X.download() - ref).max() <= 1e-9 * np.abs(ref).max()


def test_spde_elasticity_16_cubed_vs_fp64_oracle(ctx, orc):
    """The 3D elasticity statFEM at 16^3 cells (14,739 DOFs, 128 top-face
    sensors) on its production path (declared 3D grids, multigrid, tolerance
    1e-10): prior mean and C_u, the field block, the synthetic reading, and the
    posterior mean, variance and log det within 1e-8 of the FP64 oracle."""
    from statfem_amd import configs, hip
    from statfem_amd.spde_elasticity import SPDEElasticity
    cfg = configs.c3_inputs(seed=0, n=16, ns=128)
    pipe = SPDEElasticity(ctx, cfg, opts=hip.default_opts(tol=1e-10), opts_matern=hip.default_opts(tol=1e-10))
    mean, Cu = pipe.prior()
    assert pipe.stats["Rt"]["layout"] == 4 and pipe.stats["Rt"]["precond"] == 2
    y, g, u = pipe.synthesize()
    post = pipe.posterior(mean, Cu, y, field=True)
    from oracle.forward import elasticity_spde_reference
    ref = elasticity_spde_reference(cfg, y)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    errs = {"mean": rel(mean, ref["mean"]), "Cu": rel(Cu, ref["Cu"]),
            "field": rel(pipe._field.download(), ref["field"]),
            "post_mean": rel(post["mean"], ref["post_mean"]), "post_var": rel(post["var"], ref["post_var"]),
            "logdet": abs(post["logdet"] - ref["logdet"]) / abs(ref["logdet"]),
            "field_mean": rel(post["field_mean"], ref["field_mean"])}
    print("C3 16^3 errors vs FP64:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert ref["sign"] > 0
    for k, v in errs.items():
        assert v <= 1e-8, (k, v)
    # the reading is the forward model at the sampled load
    Fg = ref["F"].copy()
    Fg[2::3] += ref["w"] * g
    assert np.abs(ref["ch"].solve(Fg) - u).max() <= 1e-8 * np.abs(u).max()
