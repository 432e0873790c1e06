"""GPU parity of the statistical stages against the CPU FP64 oracle and the
reference goldens: forcing covariance, forward Gaussian model, KOH
log-likelihood, conditioned moments, dense GP and RW-MH chains."""
import os

import numpy as np
import pytest

from statfem_amd import mesh as M

from .conftest import GOLDEN
from .test_oracle import _gp_fixture, laplace_rhs_infer_problem

pytestmark = pytest.mark.gpu


def test_covforcing_dense_and_separable(ctx, orc):
    from statfem_amd import hip
    m = M.grid2d(20, 13)
    s = hip.SystemSolveHip(ctx, m)
    cv = orc.body_force(m, fconst=1.0)
    rng = np.random.default_rng(6)
    W = rng.normal(size=(m.nnodes, 7))
    ref = orc.covforcing_apply(m.coords, cv, 0.1, 0.2, W)
    Vd = s.covforcing_apply(cv, 0.1, 0.2, W, mode=1).download()
    assert np.abs(Vd - ref).max() <= 1e-13 * np.abs(ref).max()
    s.set_grid()
    Vs = s.covforcing_apply(cv, 0.1, 0.2, W, mode=2).download()
    assert np.abs(Vs - ref).max() <= 1e-13 * np.abs(ref).max()
    m3 = M.grid3d(6, 5, 4)
    s3 = hip.SystemSolveHip(ctx, m3)
    s3.set_grid()
    cv3 = orc.body_force(m3, fconst=1.0)
    W3 = rng.normal(size=(m3.nnodes, 5))
    ref3 = orc.covforcing_apply(m3.coords, cv3, 0.3, 0.25, W3)
    V3 = s3.covforcing_apply(cv3, 0.3, 0.25, W3, mode=2).download()
    assert np.abs(V3 - ref3).max() <= 1e-13 * np.abs(ref3).max()


@pytest.mark.parametrize("shape", [(1000, 37, 41), (70001, 256, 256), (33, 1, 5)])
def test_gram_mfma(ctx, shape):
    from statfem_amd import hip
    N, m, n = shape
    rng = np.random.default_rng(7)
    X = rng.normal(size=(N, m))
    V = rng.normal(size=(N, n))
    C = hip.gram(ctx, ctx.to_device(X), ctx.to_device(V))
    ref = X.T @ V
    assert np.abs(C - ref).max() <= 1e-12 * np.sqrt(N) * np.abs(ref).max()


def test_laplace_rhs_infer_golden_on_gpu(ctx, orc):
    """The reference 1D statFEM run end to end on the GPU path: LINE2 assembly,
    row-zero BCs, forward Gaussian model with P = I (33 transposed solves),
    Woodbury posterior mean at the recorded hyperparameters."""
    from statfem_amd import hip
    p = laplace_rhs_infer_problem(orc)
    m = p["mesh"]
    s = hip.SystemSolveHip(ctx, m)
    s.assemble(hip.POISSON, 1.0, nq1d=3)
    s.assemble_rhs(nq1d=3, fconst=np.pi ** 2 / 5)
    s.applyConstraints(p["bc"], np.zeros(len(p["bc"])))
    trip = np.loadtxt(os.path.join(GOLDEN, "laplaceRHSinfer", "phiMat.dat"))
    cols = trip[:, 1].astype(np.int32).reshape(-1, 1)
    vals = trip[:, 2].reshape(-1, 1)
    fg = hip.ForwardGaussRHSHip(s, cols, vals, reg_eps=p["reg"])
    mean, Cu, st = fg.updateSolution(p["sf"], p["lf"], want_field=True,
                                     opts=hip.default_opts(tol=1e-13))
    # oracle forward model (reference algorithm: transposed LU solves, dense C_f)
    lu = orc.LU(p["rp"], p["ci"], p["va"])
    R = p["P"] @ lu.solve(np.eye(m.nnodes), trans=True).T
    Cu_ref = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], R) \
        + (p["sf"] * p["reg"]) ** 2 * p["P"] @ p["P"].T
    mean_ref = R @ p["F"]
    assert np.abs(mean - mean_ref).max() <= 1e-10 * np.abs(mean_ref).max()
    assert np.abs(Cu - Cu_ref).max() <= 1e-9 * np.abs(Cu_ref).max()
    rho, sd, ld = p["hp"]
    lik = hip.LikelihoodKOHHip(ctx, p["xs"], Cu, mean, p["Y"], p["eps"])
    mfield = lik.giveFEconditionedDataWoodbury(rho, sd, ld, fg.field, fg.ubar).download()
    assert np.abs(mfield - p["ref"]).max() < 1e-6          # golden, printed to 6 digits
    Ks = orc.mismatch_cov(p["xs"], sd, ld, p["eps"])
    CuFull = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"],
                                    lu.solve(np.eye(m.nnodes), trans=True).T) \
        + (p["sf"] * p["reg"]) ** 2 * np.eye(m.nnodes)
    ubar = lu.solve(p["F"])
    mref, _ = orc.post_woodbury(CuFull, ubar, p["P"], Ks, p["Y"], rho, want_cov=False)
    assert np.abs(mfield - mref).max() <= 1e-8 * np.abs(mref).max()


def _c1_problem(orc, seed=0):
    from statfem_amd import configs
    c = configs.c1_inputs(seed)
    m = c["mesh"]
    F = orc.body_force(m, nq1d=1, fconst=0.0, fnodal=c["f_true"])
    rp, ci, va, F = orc.assemble(m, orc.POISSON, 1.0, nq1d=1, bc_dofs=c["bc"], F=F)
    lu = orc.LU(rp, ci, va)
    u = lu.solve(F)
    P = M.dense_projection(m.nnodes, c["cols"], c["vals"])
    Y = c["rho"] * (P @ u)[None, :] + c["d"] + c["e"]
    return c, m, lu, P, Y


def test_c1_statfem_end_to_end(ctx, orc):
    """Config 1 (64x64 P1, 32 sensors): prior mean/covariance at the sensors,
    log-likelihood and conditioned moments, GPU vs oracle at 1e-8."""
    from statfem_amd import hip
    c, m, lu, P, Y = _c1_problem(orc)
    s = hip.SystemSolveHip(ctx, m)
    s.assemble(hip.POISSON, 1.0, nq1d=1)
    s.assemble_rhs(nq1d=1, fconst=1.0)
    s.applyConstraints(c["bc"])
    s.set_grid()
    fg = hip.ForwardGaussRHSHip(s, c["cols"], c["vals"], reg_eps=c["reg"])
    mean, Cu, st = fg.updateSolution(c["sigma_f"], c["ell_f"], opts=hip.default_opts(tol=1e-12))
    # oracle: reference form (mean force f = 1, dense C_f over all nodes)
    rp, ci, va, F1 = orc.assemble(m, orc.POISSON, 1.0, nq1d=1, bc_dofs=c["bc"],
                                  F=orc.body_force(m, nq1d=1, fconst=1.0))
    R = orc.LU(rp, ci, va).solve(P.T.copy(), trans=True).T.copy()
    cv = orc.body_force(m, nq1d=1, fconst=1.0)
    cv[c["bc"]] = 0.0
    Cu_ref = orc.covforcing_project(m.coords, cv, c["sigma_f"], c["ell_f"], R) \
        + (c["sigma_f"] * c["reg"]) ** 2 * P @ P.T
    mean_ref = R @ F1
    assert np.abs(mean - mean_ref).max() <= 1e-9 * np.abs(mean_ref).max()
    assert np.abs(Cu - Cu_ref).max() <= 1e-8 * np.abs(Cu_ref).max()
    lik = hip.LikelihoodKOHHip(ctx, c["xs"], Cu_ref, mean_ref, Y, c["sigma_e"])
    thetas = np.array([[1.0, 0.05, 0.3], [0.8, 0.02, 0.1], [1.3, 0.1, 0.5]])
    ll, stt = lik.giveLogLikelihood(thetas)
    for t, l in zip(thetas, ll):
        Ks = orc.mismatch_cov(c["xs"], t[1], t[2], c["sigma_e"])
        ref = orc.loglik(Cu_ref, mean_ref, Y, Ks, t[0])
        assert abs(l - ref) <= 1e-10 * abs(ref)
    assert np.all(stt == 0)
    mp, cp = lik.giveProjFEconditionedData(1.0, 0.05, 0.3)
    Ks = orc.mismatch_cov(c["xs"], 0.05, 0.3, c["sigma_e"])
    mref, cref = orc.post_proj(Cu_ref, mean_ref, Y, Ks, 1.0)
    assert np.abs(mp - mref).max() <= 1e-8 * np.abs(mref).max()
    assert np.abs(cp - cref).max() <= 1e-8 * np.abs(cref).max()


def test_loglik_not_spd_status(ctx, orc):
    from statfem_amd import hip
    xs = np.array([[0.2, 0.2], [0.2, 0.2], [0.7, 0.1]])   # duplicate sensors
    Cu = np.zeros((3, 3))
    lik = hip.LikelihoodKOHHip(ctx, xs, Cu, np.zeros(3), np.ones((2, 3)), 0.0)
    ll, st = lik.giveLogLikelihood([[1.0, 0.1, 0.2], [1.0, 0.1, 0.2]])
    assert np.all(st == -4) and np.all(np.isinf(ll))


def test_gp_golden_on_gpu(ctx):
    from statfem_amd import hip
    hp, X, y, Xp, mean, var = _gp_fixture()
    ll, mu, v = hip.gaussian_process(ctx, X, y, float(hp["gaussianProcess.sigma"]),
                                     float(hp["gaussianProcess.length"]), float(hp["gaussianProcess.eps"]), Xp)
    assert abs(ll - (-4.837171259)) < 1e-9
    assert np.abs(mu - mean).max() < 1e-8 and np.abs(v - var).max() < 1e-8


def test_mcmc_chains_match_oracle(ctx, orc):
    """Same Philox streams and the same target: accept decisions and states
    of every chain agree with the oracle's sequential sampler."""
    from statfem_amd import hip
    c, m, lu, P, Y = _c1_problem(orc)
    ns = len(c["xs"])
    rng = np.random.default_rng(8)
    A = rng.normal(size=(ns, ns)) * 0.01
    Cu = A @ A.T + 1e-6 * np.eye(ns)
    mean = P @ lu.solve(orc.body_force(m, nq1d=1, fconst=1.0))
    lik = hip.LikelihoodKOHHip(ctx, c["xs"], Cu, mean, Y, c["sigma_e"])
    chains, steps = 4, 150
    th0 = np.log(np.array([[1.0, 0.05, 0.3]] * chains)) + rng.normal(size=(chains, 3)) * 0.05
    prop = np.array([0.02, 0.1, 0.1])
    for mode in (0, 2):
        samples, acc = lik.mcmc(th0, prop, steps, seed=11, prior_mode=mode,
                                prior_mu=np.log([1.0, 0.05, 0.3]), prior_var=[0.1, 0.5, 0.5])
        for ch in range(chains):
            ref, racc = orc.mcmc_koh(Cu, mean, Y, c["xs"], c["sigma_e"], th0[ch], prop, steps, seed=11,
                                     chain=ch, prior_mode=mode, prior_mu=np.log([1.0, 0.05, 0.3]),
                                     prior_var=[0.1, 0.5, 0.5])
            assert acc[ch] == racc
            assert np.abs(samples[ch, :, :3] - ref[:, :3]).max() < 1e-12
            assert np.abs(samples[ch, :, 3] - ref[:, 3]).max() < 1e-8 * np.abs(ref[:, 3]).max()


def test_typical_readings_match_oracle(ctx, orc):
    from statfem_amd import hip
    rng = np.random.default_rng(10)
    ns = 40
    xs = rng.uniform(size=(ns, 2))
    A = rng.normal(size=(ns, ns)) * 0.05
    Cu = A @ A.T + 1e-6 * np.eye(ns)
    mean = rng.normal(size=ns)
    lik = hip.LikelihoodKOHHip(ctx, xs, Cu, mean, np.zeros((1, ns)), 0.01)
    Yg = lik.giveTypicalReadings(1.2, 0.05, 0.2, 17, seed=123)
    Yo = orc.typical_readings(Cu, mean, xs, 0.01, 1.2, 0.05, 0.2, 17, seed=123)
    assert np.abs(Yg - Yo).max() < 1e-12 * np.abs(Yo).max()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_precision_operator_matches_oracle(ctx, orc, dim):
    from statfem_amd import hip
    m = {1: M.line1d(60), 2: M.grid2d(14, 11), 3: M.grid3d(5, 4, 6)}[dim]
    nu = {1: 1.5, 2: 1.0, 3: 0.5}[dim]
    kap = np.sqrt(2 * nu) / 0.2
    s = hip.SystemSolveHip(ctx, m)
    s.assemble(hip.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mass = s.lumped_mass(2)
    rng = np.random.default_rng(11)
    X = rng.normal(size=(m.nnodes, 5))
    Yg = s.precision_apply(mass, 0.8, kap, nu, X).download()
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    Yo = orc.precision_apply(rp, ci, va, orc.body_force(m, nq1d=2, fconst=1.0), 0.8, kap, nu, dim, X)
    assert np.abs(Yg - Yo).max() <= 1e-12 * np.abs(Yo).max()
    assert abs(hip.matern_tau(0.8, kap, nu, dim) - orc.matern_tau(0.8, kap, nu, dim)) < 1e-15
    Ys = s.spmv(X).download()
    assert np.abs(Ys - orc.spmv(rp, ci, va, X)).max() <= 1e-13 * np.abs(Ys).max()
    with pytest.raises(hip.SfhError):
        s.precision_apply(mass, 0.8, kap, nu + 0.3, X)   # fractional beta without its factors


@pytest.mark.parametrize("dim,nu", [(1, 1.0), (2, 0.5), (2, 1.5)])
def test_fractional_precision_operator_matches_oracle(ctx, orc, dim, nu):
    """P_l, P_r, P_l^T and Q_v = P_l^T Q P_l applied to a block on the GPU
    against the oracle's dense matrices, with the degree-4 rational factors."""
    from statfem_amd import hip, rational
    m = {1: M.line1d(60), 2: M.grid2d(14, 11)}[dim]
    kap = np.sqrt(2 * nu) / 0.2
    beta = (nu + dim / 2) / 2
    top, bot, _ = rational.brasil_factors(beta - max(1.0, np.floor(beta)), 4)
    s = hip.SystemSolveHip(ctx, m)
    s.assemble(hip.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mass = s.lumped_mass(2)
    X = np.random.default_rng(12).normal(size=(m.nnodes, 5))
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mh = orc.body_force(m, nq1d=2, fconst=1.0)
    for roots in (top, bot):
        F = orc.gpq_frac_factor(rp, ci, va, mh, roots)
        for tr, ref in ((False, F @ X), (True, F.T @ X)):
            got = s.frac_apply(mass, roots, X, transpose=tr).download()
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (tr, roots[0])
    Qv, _ = orc.gpq_precision_frac_dense(rp, ci, va, mh, 0.8, kap, nu, dim, top, bot)
    ref = Qv @ X
    got = s.precision_apply(mass, 0.8, kap, nu, X, bot=bot).download()
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("ns,no", [(20, 3), (48, 16), (64, 30), (100, 4), (150, 9), (177, 31), (200, 8), (215, 9),
                                   (240, 6)])
def test_loglik_sizes_match_oracle(ctx, orc, ns, no):
    """Every factorisation path of the batched log-likelihood (the blocked MFMA
    Cholesky from 48 sensors while the padded sensor tiles and the data tiles
    fit 14 tiles of 16, the register-resident one up to 224 augmented rows
    otherwise, the work-matrix form beyond) against the oracle, with its
    log-determinant, over several hyperparameter sets; sensor counts off and
    on the tile size, one and two data tiles."""
    from statfem_amd import hip
    rng = np.random.default_rng(ns)
    xs = rng.uniform(size=(ns, 2))
    A = rng.normal(size=(ns, ns)) * 0.05
    Cu = A @ A.T + 1e-4 * np.eye(ns)
    mean = rng.normal(size=ns)
    Y = rng.normal(size=(no, ns))
    lik = hip.LikelihoodKOHHip(ctx, xs, Cu, mean, Y, 0.01)
    thetas = np.array([[1.0, 0.05, 0.3], [0.7, 0.2, 0.1], [1.4, 0.01, 0.6]])
    ll, st = lik.giveLogLikelihood(thetas)
    ld, _ = lik.giveLogDeterminant(thetas)
    assert np.all(st == 0)
    for t, l, d in zip(thetas, ll, ld):
        Ks = orc.mismatch_cov(xs, t[1], t[2], 0.01)
        ref = orc.loglik(Cu, mean, Y, Ks, t[0])
        assert abs(l - ref) <= 1e-10 * abs(ref)
        assert abs(d - np.linalg.slogdet(Ks + t[0] ** 2 * Cu)[1]) <= 1e-10 * abs(d)
    bad = hip.LikelihoodKOHHip(ctx, xs, -np.eye(ns), mean, Y, 0.0)
    llb, stb = bad.giveLogLikelihood([[1.0, 0.0, 0.3]])
    assert stb[0] == -4 and np.isinf(llb[0])
