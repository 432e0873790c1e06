"""Learned surrogate: GPU forward/backward against the FP64 oracle MLP,
training, and delayed-acceptance chains that must reproduce the plain
chain's posterior."""
import numpy as np
import pytest

from .test_gpu_stats import _c1_problem

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle(ctx):
    from oracle import mlp
    from statfem_amd import hip
    net = hip.SurrogateMLP(ctx, seed=3)
    rng = np.random.default_rng(13)
    th = rng.normal(size=(300, 3))
    y = np.sin(th[:, 0]) + th[:, 1] * th[:, 2]
    norm = np.array([0.1, -0.2, 0.05, 1.1, 0.9, 1.3, 0.2, 1.7])
    net.set_normalization(norm)
    g, loss = net.grad(th, y)
    _, bf, _ = net.params()
    p = mlp.bf16_to_float(bf)
    X = (th - norm[:3]) / norm[3:6]
    t = (y - norm[6]) / norm[7]
    loss_ref, g_ref = mlp.loss_grad(p, X, t)
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert np.linalg.norm(g - g_ref) < 1e-4 * np.linalg.norm(g_ref)
    pred = net.predict(th)
    assert np.abs(pred - (mlp.forward(p, X)[0] * norm[7] + norm[6])).max() < 1e-4 * norm[7]


def test_mlp_training_fits_a_smooth_surface(ctx):
    from statfem_amd import hip
    net = hip.SurrogateMLP(ctx, seed=4)
    rng = np.random.default_rng(14)
    th = rng.uniform(-1, 1, size=(4096, 3))
    f = lambda t: -3 * t[:, 0] ** 2 - (t[:, 1] - 0.3 * t[:, 2]) ** 2 + 0.5 * t[:, 2]
    loss = net.train(th, f(th), steps=1500, lr=3e-3)
    assert loss < 2e-3
    tv = rng.uniform(-0.9, 0.9, size=(500, 3))
    assert np.sqrt(np.mean((net.predict(tv) - f(tv)) ** 2)) < 0.05 * f(tv).std()


def test_delayed_acceptance_reproduces_plain_posterior(ctx, orc):
    """The exact second stage keeps the posterior whatever the surrogate's
    error: DA and plain chains agree on means within their Monte-Carlo
    standard errors and on standard deviations."""
    from statfem_amd import diagnostics as D
    from statfem_amd import hip
    c, m, lu, P, Y = _c1_problem(orc)
    ns = len(c["xs"])
    rng = np.random.default_rng(15)
    A = rng.normal(size=(ns, ns)) * 0.01
    Cu = A @ A.T + 1e-6 * np.eye(ns)
    mean = P @ lu.solve(orc.body_force(m, nq1d=1, fconst=1.0))
    lik = hip.LikelihoodKOHHip(ctx, c["xs"], Cu, mean, Y, c["sigma_e"])
    pm, pv = np.log([1.0, 0.05, 0.3]), np.array([0.5, 0.5, 0.5])
    kw = dict(prior_mode=2, prior_mu=pm, prior_var=pv)
    chains = 64
    th0 = np.tile(pm, (chains, 1)) + rng.normal(size=(chains, 3)) * 0.05
    pilot, _ = lik.mcmc(th0, [0.02, 0.02, 0.02], 2000, seed=7, **kw)
    sd = pilot[:, 1000:, :3].reshape(-1, 3).std(0)
    prop = 1.2 * sd
    start = pilot[:, -1, :3].copy()
    plain, _ = lik.mcmc(start, prop, 4000, seed=1, **kw)
    plain = plain[:, 500:, :3]
    post = plain.reshape(-1, 3)
    tr = post.mean(0) + rng.normal(size=(2048, 3)) * 2.5 * post.std(0)
    ll, st = lik.giveLogLikelihood(np.exp(tr))
    ok = st == 0
    net = hip.SurrogateMLP(ctx, seed=5)
    net.train(tr[ok], ll[ok], steps=2000, lr=3e-3)
    da, dacc, sacc = hip.mcmc_da(lik, net, start, prop, 1000, 8, seed=2, **kw)
    da = da[:, 100:, :3]
    dpost = da.reshape(-1, 3)
    se = np.sqrt(D.mcse_mean(plain) ** 2 + D.mcse_mean(da) ** 2)
    assert np.all(np.abs(dpost.mean(0) - post.mean(0)) < 4.5 * se)
    assert np.all(np.abs(dpost.std(0) / post.std(0) - 1) < 0.1)
    assert np.all(D.ess_all(da) / da.shape[1] > D.ess_all(plain) / plain.shape[1])
    assert dacc.mean() > 0.3
