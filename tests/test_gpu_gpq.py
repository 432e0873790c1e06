"""Sparse-precision GP (GaussianProcessQ with the SPDE precision, integer and
fractional beta) on the GPU against the oracle's dense restatement of the
reference algorithm and the exact Gaussian densities."""
import numpy as np
import pytest

from statfem_amd import mesh as M

pytestmark = pytest.mark.gpu


def _case(ctx, orc, dim, declare):
    from statfem_amd import hip
    from statfem_amd.gpq import GaussianProcessQHip
    m = M.grid2d(20, 15) if dim == 2 else M.line1d(70)
    rng = np.random.default_rng(dim + 10 * declare)
    s = hip.SystemSolveHip(ctx, m)
    if declare:
        s.set_grid()
    ns, no = 13, 3
    if dim == 2:
        xs = rng.uniform(0.05, 0.95, size=(ns, 2))
        cols, vals = M.sensor_rows(m, xs)
    else:
        cols = rng.choice(m.nnodes, ns, replace=False).astype(np.int32).reshape(ns, 1)
        vals = np.ones((ns, 1))
    Y = 0.4 * rng.standard_normal((ns, no))
    gp = GaussianProcessQHip(s, cols, vals, Y, nq1d=2)
    P = M.dense_projection(m.nnodes, cols, vals)
    return m, gp, P, Y


def _oracle_q(orc, m, sigma, kappa, nu):
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kappa, 1.0, nq1d=2, factor=1.0 / kappa ** 2)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    return orc.gpq_precision_dense(rp, ci, va, mass, sigma, kappa, nu, m.dim)


def _joint_loglik(orc, m, P, Y, sigma, kappa, nu, eps):
    """log p(Y) as the density of the stacked readings (covariance
    J (x) G + eps^2 I) with G = P Q^-1 P^T from Q^-1 = (1/c) A^-1 (M A^-1)^(2 beta - 1)
    by dense solves with A (condition ~ cond(A), where Q's is ~ cond(A)^(2 beta))."""
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kappa, 1.0, nq1d=2, factor=1.0 / kappa ** 2)
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] += va[rp[i]:rp[i + 1]]
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    beta = int(round((nu + m.dim / 2) / 2))
    c = orc.matern_tau(sigma, kappa, nu, m.dim) ** 2 * kappa ** (4 * beta)
    Ai = np.linalg.inv(A)
    X = Ai.copy()
    for _ in range(2 * beta - 1):
        X = Ai @ (mass[:, None] * X)
    G = P @ X @ P.T / c
    ns, no = Y.shape
    K = np.kron(np.ones((no, no)), G) + eps * eps * np.eye(ns * no)
    yv = Y.T.reshape(-1)
    _, ld = np.linalg.slogdet(K)
    return -0.5 * (yv @ np.linalg.solve(K, yv) + ld + ns * no * np.log(2 * np.pi)), beta


@pytest.mark.parametrize("dim,declare", [(1, 0), (2, 0), (2, 1)])
def test_loglik_matches_oracle(ctx, orc, dim, declare):
    """Three (sigma, kappa, nu) groups, two of them beta = 2, each with several
    noise levels, in one call; 2D both on CSR and on the declared grid
    (multigrid)."""
    from statfem_amd import hip
    m, gp, P, Y = _case(ctx, orc, dim, declare)
    nus = (1.5, 3.5) if dim == 1 else (1.0, 3.0)
    sets = []
    for sigma, ell, nu in [(0.8, 0.3, nus[0]), (1.3, 0.15, nus[0]), (0.6, 0.4, nus[1])]:
        kap = np.sqrt(2 * nu) / ell
        for eps in (0.03, 0.1, 0.4):
            sets.append([sigma, kap, nu, eps])
    ll, st, stats = gp.giveLogLikelihood(sets, hip.default_opts(tol=1e-12))
    assert np.all(st == 0)
    if declare:
        assert stats["precond"] == 2
    for row, l in zip(sets, ll):
        # the reference algorithm: two Cholesky log-determinants of Q and
        # eps^2 Q + n_o P^T P, whose difference loses ~cond(Q) eps64 (Q's
        # condition grows as cond(A)^(2 beta): ~1e10 at beta = 2 here)
        Q = _oracle_q(orc, m, row[0], row[1], row[2])
        ref = orc.gpq_loglik(Q, P, Y, row[3])
        exact, beta = _joint_loglik(orc, m, P, Y, *row)
        assert abs(l - exact) <= 1e-9 * abs(exact), (row, l, exact)
        assert abs(l - ref) <= (1e-8 if beta == 1 else 1e-6) * abs(ref), (row, l, ref)
    gp.close()


@pytest.mark.parametrize("dim", [1, 2])
def test_posterior_matches_oracle(ctx, orc, dim):
    from statfem_amd import hip
    m, gp, P, Y = _case(ctx, orc, dim, dim == 2)
    nu = 1.0 if dim == 2 else 1.5
    hp = [0.9, np.sqrt(2 * nu) / 0.25, nu, 0.08]
    o = hip.default_opts(tol=1e-12)
    mean = gp.predictMean(hp, o)
    cov = gp.predictCovariance(hp, o)
    Q = _oracle_q(orc, m, *hp[:3])
    mref, cref = orc.gpq_predict(Q, P, Y, hp[3])
    assert np.abs(mean - mref).max() <= 1e-8 * np.abs(mref).max()
    assert np.abs(cov - cref).max() <= 1e-8 * np.abs(cref).max()
    var = gp.predictVariance(hp, o)
    assert np.abs(var - np.diag(cref)).max() <= 1e-8 * np.abs(cref).max()
    gp.close()


def test_driver_keys_and_fractional_beta(ctx, orc):
    """DriverMaxLikelihood's map: kappa given directly or through the length
    (kappa = sqrt(2 nu) / length); the C ABI refuses a fractional exponent
    without its rational factors and a non-negative pole, and the Python
    mirror supplies the factors by itself."""
    import ctypes as C
    from statfem_amd import rational
    from statfem_amd import hip
    from statfem_amd.gpq import DriverMaxLikelihoodHip
    m, gp, P, Y = _case(ctx, orc, 2, 1)
    drv = DriverMaxLikelihoodHip(gp)
    drv.setAllHps({"gaussianProcess.sigma": 0.9, "gaussianProcess.nu": 1.0, "gaussianProcess.length": 0.25,
                   "gaussianProcess.eps": 0.05})
    l1 = drv.computeLogLikelihood(hip.default_opts(tol=1e-12))
    h = drv.giveAllHps()
    h.pop("gaussianProcess.length")
    h["gaussianProcess.kappa"] = np.sqrt(2.0) / 0.25
    drv.setAllHps(h)
    l2 = drv.computeLogLikelihood(hip.default_opts(tol=1e-12))
    assert abs(l1 - l2) <= 1e-10 * abs(l1)
    drv.addOptimHps("gaussianProcess.eps")
    ll, st = drv.computeLogLikelihoods([[0.02], [0.05], [0.2]], hip.default_opts(tol=1e-12))
    assert np.all(st == 0) and abs(ll[1] - l2) <= 1e-10 * abs(l2)
    Q = _oracle_q(orc, m, 0.9, np.sqrt(2.0) / 0.25, 1.0)
    assert abs(ll[0] - orc.gpq_loglik(Q, P, Y, 0.02)) <= 1e-8 * abs(ll[0])
    t = np.array([[0.9, 5.0, 1.5, 0.05]])            # beta = 1.25 in 2D
    out, st = np.zeros(1), np.zeros(1, np.int32)
    o = hip.default_opts(tol=1e-12)
    rc = hip.lib().sfh_gpq_loglik(gp.h, 1, hip._p(t), C.byref(o), hip._p(out), hip._p(st), None)
    assert hip.STATUS[rc] == "SFH_EARG"
    top, bot, _ = rational.brasil_factors(0.25, 4)
    bad = bot.copy()
    bad[2] = 0.1
    with pytest.raises(hip.SfhError):
        gp.set_rational(0.25, top, bad)
    l_auto = gp.giveLogLikelihood(t, o)[0][0]          # factors computed on first use
    gp.set_rational(0.25, top, bot)
    l_set = gp.giveLogLikelihood(t, o)[0][0]
    assert np.isfinite(l_auto) and abs(l_auto - l_set) <= 1e-12 * abs(l_set)
    gp.close()


def _exact_fractional(orc, m, sigma, kappa, nu, top, bot):
    """The field covariance P_r Q_v^-1 P_r^T in the generalized eigenbasis of
    (A, M): (1/c) V diag(r(1/lam)^2 lam^(-2 alpha)) V^T, r the factor form."""
    from scipy.linalg import eigh
    from statfem_amd import rational
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kappa, 1.0, nq1d=2, factor=1.0 / kappa ** 2)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    lam, V = eigh(orc.spmv(rp, ci, va, np.eye(m.nnodes)), np.diag(mass))
    beta = (nu + m.dim / 2) / 2
    alpha = max(1.0, np.floor(beta))
    c = orc.matern_tau(sigma, kappa, nu, m.dim) ** 2 * kappa ** (4 * beta)
    rr = rational.factor_form(top, bot, 1.0 / lam)
    return V @ np.diag(rr ** 2 * lam ** (-2 * alpha)) @ V.T / c, (rp, ci, va, mass)


@pytest.mark.parametrize("dim,declare,nu", [(1, 0, 1.0), (2, 0, 0.5), (2, 1, 0.5), (2, 1, 1.5), (2, 1, 2.5)])
def test_fractional_beta_matches_oracle(ctx, orc, dim, declare, nu):
    """Fractional beta (2D: 0.75, 1.25, 1.75; 1D: 0.75) with the rational
    factors of the reference's degree 4: likelihood at three noise levels,
    posterior mean and covariance, on CSR and on the declared grid."""
    from statfem_amd import hip, rational
    m, gp, P, Y = _case(ctx, orc, dim, declare)
    e = gp.exponent(nu)
    assert e != 0.0
    top, bot, _ = rational.brasil_factors(e, 4)
    gp.set_rational(e, top, bot)
    sigma, kap = 0.8, np.sqrt(2 * nu) / 0.3
    o = hip.default_opts(tol=1e-12)
    sets = [[sigma, kap, nu, eps] for eps in (0.03, 0.1, 0.4)]
    ll, st, stats = gp.giveLogLikelihood(sets, o)
    assert np.all(st == 0)
    if declare:
        assert stats["precond"] == 2
    Cu, (rp, ci, va, mass) = _exact_fractional(orc, m, sigma, kap, nu, top, bot)
    Qv, Pr = orc.gpq_precision_frac_dense(rp, ci, va, mass, sigma, kap, nu, m.dim, top, bot)
    ns, no = Y.shape
    Pb = np.vstack([P] * no)
    yv = Y.T.reshape(-1)
    for row, l in zip(sets, ll):
        K = Pb @ Cu @ Pb.T + row[3] ** 2 * np.eye(ns * no)
        _, ld = np.linalg.slogdet(K)
        exact = -0.5 * (yv @ np.linalg.solve(K, yv) + ld + ns * no * np.log(2 * np.pi))
        assert abs(l - exact) <= 1e-8 * abs(exact), (row, l, exact)
        # the reference algorithm on Q_v: its two log-determinants carry the
        # condition of Q_v, cond(Q) cond(P_l)^2 (largest in 1D), so it agrees
        # to about 1e-6 only; the exact density above is the pin
        ref = orc.gpq_loglik(Qv, P, Y, row[3], Pr)
        assert abs(l - ref) <= 1e-5 * abs(ref), (row, l, ref)
    hp = [sigma, kap, nu, 0.1]
    mean = gp.predictMean(hp, o)
    cov = gp.predictCovariance(hp, o)
    K = Pb @ Cu @ Pb.T + hp[3] ** 2 * np.eye(ns * no)
    mref = Cu @ Pb.T @ np.linalg.solve(K, yv)
    cref = Cu - Cu @ Pb.T @ np.linalg.solve(K, Pb @ Cu)
    assert np.abs(mean - mref).max() <= 1e-8 * np.abs(mref).max()
    assert np.abs(cov - cref).max() <= 1e-8 * np.abs(cref).max()
    gp.close()
