"""The C++ drop-ins of the statX classes (AssemblerHip, ForwardGaussRHSHip,
CovarianceForcingHip, MismatchCovarianceHip, LikelihoodKOHHip,
GaussianProcessHip) driven by tests/cpp/statfem_pipeline.cpp like the
laplaceRHSinfer driver, against the oracle and the reference goldens; the
Woodbury field posterior covariance and GaussianProcess::predictCovariance on
the GPU; and the error statuses of the C ABI (singular system, iteration cap,
non-negative quadratic form)."""
import subprocess

import numpy as np
import pytest

from statfem_amd import mesh as M

from .conftest import GOLDEN
from .test_oracle import _gp_fixture, laplace_rhs_infer_problem

pytestmark = pytest.mark.gpu


def _read_out(path):
    out = {}
    for line in open(path):
        parts = line.split()
        out[parts[0]] = np.array([float(v) for v in parts[2:2 + int(parts[1])]])
    return out


def test_cpp_statx_pipeline(orc, tmp_path):
    from statfem_amd import build
    path = tmp_path / "pipeline.txt"
    r = subprocess.run([build.CPP_PIPELINE, GOLDEN, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    o = _read_out(path)
    p = laplace_rhs_infer_problem(orc)
    m = p["mesh"]
    N = m.nnodes
    lu = orc.LU(p["rp"], p["ci"], p["va"])
    u = lu.solve(p["F"])
    assert np.abs(o["u"] - u).max() <= 1e-10 * np.abs(u).max()
    assert np.abs(o["covvec"] - p["covvec"]).max() <= 1e-15
    Rfull = lu.solve(np.eye(N), trans=True).T
    R = p["P"] @ Rfull
    ns = R.shape[0]
    Cu = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], R) \
        + (p["sf"] * p["reg"]) ** 2 * p["P"] @ p["P"].T
    mean = R @ p["F"]
    assert np.abs(o["mean"] - mean).max() <= 1e-10 * np.abs(mean).max()
    assert np.abs(o["Cu"].reshape(ns, ns) - Cu).max() <= 1e-9 * np.abs(Cu).max()
    rho, sd, ld = p["hp"]
    Ks = orc.mismatch_cov(p["xs"], sd, ld, p["eps"])
    assert np.abs(o["Ks"].reshape(ns, ns) - Ks).max() <= 1e-15
    # kernel derivatives against central differences of K_s
    h = 1e-6
    for name, dk in (("dKs_sigma", (orc.mismatch_cov(p["xs"], sd + h, ld, p["eps"]) -
                                    orc.mismatch_cov(p["xs"], sd - h, ld, p["eps"])) / (2 * h)),
                     ("dKs_length", (orc.mismatch_cov(p["xs"], sd, ld + h, p["eps"]) -
                                     orc.mismatch_cov(p["xs"], sd, ld - h, p["eps"])) / (2 * h))):
        assert np.abs(o[name].reshape(ns, ns) - dk).max() <= 1e-6 * np.abs(dk).max(), name
    Cu_gpu = o["Cu"].reshape(ns, ns)
    mean_gpu = o["mean"]
    ll = orc.loglik(Cu_gpu, mean_gpu, p["Y"], Ks, rho)
    assert abs(o["logL"][0] - ll) <= 1e-10 * abs(ll)
    for b, (rb, sb, lb) in enumerate(((rho, sd, ld), (1.1, sd, ld), (rho, sd, 0.2))):
        ref = orc.loglik(Cu_gpu, mean_gpu, p["Y"], orc.mismatch_cov(p["xs"], sb, lb, p["eps"]), rb)
        assert abs(o["logL_batch"][b] - ref) <= 1e-10 * abs(ref)
    pm, pc = orc.post_proj(Cu_gpu, mean_gpu, p["Y"], Ks, rho)
    assert np.abs(o["proj_mean"] - pm).max() <= 1e-8 * np.abs(pm).max()
    assert np.abs(o["proj_cov"].reshape(ns, ns) - pc).max() <= 1e-8 * np.abs(pc).max()
    CuFull = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], Rfull) \
        + (p["sf"] * p["reg"]) ** 2 * np.eye(N)
    wm, wc = orc.post_woodbury(CuFull, u, p["P"], Ks, p["Y"], rho)
    assert np.abs(o["wood_mean"] - wm).max() <= 1e-8 * np.abs(wm).max()
    assert np.abs(o["wood_cov"].reshape(N, N) - wc).max() <= 1e-8 * np.abs(wc).max()
    assert np.abs(o["wood_mean"] - p["ref"]).max() < 1e-6        # golden, printed to 6 digits
    hp, X, y, Xp, _, var = _gp_fixture()
    gc = orc.gp_cov(X, float(hp["gaussianProcess.sigma"]), float(hp["gaussianProcess.length"]),
                    float(hp["gaussianProcess.eps"]), Xp, len(y))
    assert np.abs(o["gp_cov"].reshape(len(Xp), len(Xp)) - gc).max() <= 1e-10


def test_woodbury_field_covariance_2d(ctx, orc):
    """Posterior field covariance C_u - (C_u P^T) S^-1 (P C_u) and its
    diagonal on a 24 x 24 P1 grid (625 nodes, 16 sensors): the GPU's prior
    (updateSolution with P = I) and posterior against the oracle's Woodbury
    form with direct solves."""
    from statfem_amd import configs, hip
    c = configs.c1_inputs(seed=2, n=24, ns=16, no=6)
    m, bc = c["mesh"], c["bc"]
    N = m.nnodes
    s = hip.SystemSolveHip(ctx, m)
    s.set_grid()
    s.assemble(hip.POISSON, 1.0, nq1d=1)
    s.assemble_rhs(nq1d=1, fconst=1.0)
    s.applyConstraints(bc)
    o = hip.default_opts(tol=1e-12)
    fg = hip.ForwardGaussRHSHip(s, c["cols"], c["vals"], reg_eps=c["reg"])
    mean, Cu, _ = fg.updateSolution(c["sigma_f"], c["ell_f"], want_field=True, opts=o)
    full = hip.ForwardGaussRHSHip(s, np.arange(N, dtype=np.int32)[:, None], np.ones((N, 1)), reg_eps=c["reg"])
    _, CuFull_gpu, _ = full.updateSolution(c["sigma_f"], c["ell_f"], opts=o)
    # oracle prior on the whole field (P = I)
    F = orc.body_force(m, nq1d=1, fconst=1.0)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, 1.0, nq1d=1, bc_dofs=bc, F=F)
    lu = orc.LU(rp, ci, va)
    Rfull = lu.solve(np.eye(N), trans=True).T
    cv = orc.body_force(m, nq1d=1, fconst=1.0)
    cv[bc] = 0.0
    reg2 = (c["sigma_f"] * c["reg"]) ** 2
    Cf = orc.covforcing_apply(m.coords, cv, c["sigma_f"], c["ell_f"], np.eye(N))
    CuFull = Rfull @ Cf @ Rfull.T + reg2 * np.eye(N)
    assert np.abs(CuFull_gpu - CuFull).max() <= 1e-9 * np.abs(CuFull).max()
    P = M.dense_projection(N, c["cols"], c["vals"])
    rng = np.random.default_rng(4)
    Y = mean[None, :] + 1e-3 * rng.standard_normal((6, len(mean)))
    lik = hip.LikelihoodKOHHip(ctx, c["xs"], Cu, mean, Y, c["sigma_e"])
    rho, sd, ld = 1.1, 0.04, 0.25
    post = lik.giveFEconditionedCovarianceWoodbury(rho, sd, ld, fg.field, CuFull_gpu, full=True).download()
    var = lik.giveFEconditionedCovarianceWoodbury(rho, sd, ld, fg.field, np.diag(CuFull_gpu).copy(),
                                                  full=False).download()
    mpost = lik.giveFEconditionedDataWoodbury(rho, sd, ld, fg.field, fg.ubar).download()
    Ks = orc.mismatch_cov(c["xs"], sd, ld, c["sigma_e"])
    mref, cref = orc.post_woodbury(CuFull, lu.solve(F), P, Ks, Y, rho)
    assert np.abs(post - cref).max() <= 1e-8 * np.abs(cref).max()
    assert np.abs(var - np.diag(cref)).max() <= 1e-8 * np.abs(np.diag(cref)).max()
    assert np.abs(mpost - mref).max() <= 1e-8 * np.abs(mref).max()
    # the posterior variance never exceeds the prior's; on the boundary both are the ridge (sigma_f reg)^2
    assert np.all(var <= np.diag(CuFull_gpu) * (1 + 1e-12))
    assert np.abs(var[bc] - reg2).max() <= 1e-12 * reg2


def test_gp_predict_covariance_gpu(ctx, orc):
    from statfem_amd import hip
    hp, X, y, Xp, mean, var = _gp_fixture()
    sg, lg, eg = (float(hp[f"gaussianProcess.{k}"]) for k in ("sigma", "length", "eps"))
    cov = hip.gp_predict_covariance(ctx, X, y, sg, lg, eg, Xp)
    ref = orc.gp_cov(X, sg, lg, eg, Xp, len(y))
    assert np.abs(cov - ref).max() <= 1e-10
    assert np.abs(np.diag(cov) - var).max() < 1e-8               # golden variances
    rng = np.random.default_rng(8)
    X2 = rng.uniform(size=(40, 2))
    y2 = np.sin(3 * X2[:, 0]) * X2[:, 1]
    Xp2 = rng.uniform(size=(70, 2))
    c2 = hip.gp_predict_covariance(ctx, X2, y2, 0.7, 0.35, 1e-2, Xp2)
    r2 = orc.gp_cov(X2, 0.7, 0.35, 1e-2, Xp2, 40)
    assert np.abs(c2 - r2).max() <= 1e-10 * np.abs(r2).max()


@pytest.mark.parametrize("grid", [False, True], ids=["csr", "stencil"])
def test_singular_system_status(ctx, grid):
    """An operator without Dirichlet rows (pure Neumann Poisson) and a
    right-hand side with a component in its null space: the residual stops
    falling and the solve reports SFH_ESINGULAR, as the reference's SparseLU
    reports the singular factorisation (corlib/SystemSolveEigenSparse.hpp:303-305)."""
    from statfem_amd import hip
    m = M.grid2d(12)
    s = hip.SystemSolveHip(ctx, m)
    if grid:
        s.set_grid()
    s.assemble(hip.POISSON, 1.0, nq1d=2)
    B = np.random.default_rng(1).standard_normal((m.nnodes, 4)) + 1.0
    with pytest.raises(hip.SfhError) as e:
        s.solve(B, opts=hip.default_opts(tol=1e-10, max_outer=60))
    assert e.value.status == -3, str(e.value)
    # the same operator with its boundary constrained solves
    s.applyConstraints(M.dirichlet_all_boundary(m))
    _, st = s.solve(B, opts=hip.default_opts(tol=1e-10))
    assert st["converged"] == 1


def test_iteration_cap_status(ctx):
    from statfem_amd import hip
    m = M.grid2d(20)
    s = hip.SystemSolveHip(ctx, m)
    s.assemble(hip.POISSON, 1.0, nq1d=2)
    s.applyConstraints(M.dirichlet_all_boundary(m))
    B = np.random.default_rng(2).standard_normal((m.nnodes, 3))
    with pytest.raises(hip.SfhError) as e:
        s.solve(B, opts=hip.default_opts(tol=1e-12, max_inner=3, layout=1, precond=1))
    assert e.value.status == -6, str(e.value)


def test_quadratic_form_status(ctx, orc):
    """A reading equal to rho times the projected mean gives a zero quadratic
    form, which the reference rejects (LikelihoodKOH.ipp:113): SFH_EQUADFORM
    for that set, the others evaluate."""
    from statfem_amd import hip
    rng = np.random.default_rng(3)
    ns = 12
    A = rng.standard_normal((ns, ns))
    Cu = A @ A.T / ns * 1e-3
    mean = rng.standard_normal(ns)
    xs = rng.uniform(size=(ns, 2))
    Y = mean[None, :] + 0.01 * rng.standard_normal((4, ns))
    Y[2] = mean                                    # rho = 1: d = y - rho mean = 0
    lik = hip.LikelihoodKOHHip(ctx, xs, Cu, mean, Y, 1e-3)
    ll, st = lik.giveLogLikelihood([[1.0, 0.05, 0.3], [1.2, 0.05, 0.3]])
    assert st[0] == -5 and st[1] == 0
    ref = orc.loglik(Cu, mean, Y, orc.mismatch_cov(xs, 0.05, 0.3, 1e-3), 1.2)
    assert abs(ll[1] - ref) <= 1e-10 * abs(ref)
