"""Conductivity (LHS) uncertainty on the GPU: ForwardGaussLHS with the
laplaceInverse solve function, its log target and its MCMC, against the
oracle and the reference's recorded chain."""
import numpy as np
import pytest

from . import lhs_deck

pytestmark = pytest.mark.gpu


def _fg(ctx, d, copies, opts=None):
    from statfem_amd.lhs import ForwardGaussLHSHip
    fg = ForwardGaussLHSHip(ctx, d["mesh"], d["bc"], d["anchors"], d["gp"], d["cols"], d["vals"], copies=copies,
                            nq1d=3, opts=opts)
    hp = d["hp"]
    fg.set_data(d["Y"], d["xs"], rho=hp["rho"], mis_sigma=hp["mismatch.sigma"], mis_ell=hp["mismatch.length"],
                prior_mu=d["pm"], prior_var=d["pv"])
    return fg


def _tight():
    from statfem_amd import hip
    return hip.default_opts(tol=1e-12)


def test_fixed_mean_targets_reproduce_recorded_chain(ctx, orc):
    """Reference semantics (the FE solution of the map's conductivities, kept for
    every evaluation): the log targets of the recorded states equal the
    recorded values to print precision and the oracle's to 1e-10."""
    d = lhs_deck.load()
    fg = _fg(ctx, d, 1, _tight())
    fg.updateSolution(np.r_[np.log(d["w0"]), np.log(d["hp"]["mismatch.sensorEps"])])
    fg.fix_mean(0)
    th = np.log(d["golden"][:, :6])
    y, st, _ = fg.log_target(th, resolve=False)
    assert (st == 0).all()
    assert np.abs(y - d["golden"][:, 6]).max() < 1e-3
    P = lhs_deck.projection(d, d["mesh"].nnodes)
    u, _ = orc.lhs_solution(d["mesh"], d["bc"], d["anchors"], d["w0"], *d["gp"], nq1d=3)
    yo = np.array([lhs_deck.oracle_target(orc, d, t, P @ u) for t in th])
    assert np.abs(y - yo).max() <= 1e-10 * np.abs(yo).max()
    fg.close()


def _sets(d, n, seed):
    rng = np.random.default_rng(seed)
    return d["init"] + 0.3 * rng.standard_normal((n, 6)) * np.r_[np.ones(5), 0.5]


def test_update_solution_matches_oracle(ctx, orc):
    """updateSolution for 6 copies at once: each copy's projected mean is the
    oracle's direct solve at its own interpolated conductivity."""
    d = lhs_deck.load()
    fg = _fg(ctx, d, 6, _tight())
    th = _sets(d, 6, 1)
    mean, st = fg.updateSolution(th)
    P = lhs_deck.projection(d, d["mesh"].nnodes)
    for c in range(6):
        u, _ = orc.lhs_solution(d["mesh"], d["bc"], d["anchors"], np.exp(th[c, :5]), *d["gp"], nq1d=3)
        ref = P @ u
        assert np.abs(mean[c] - ref).max() <= 1e-9 * np.abs(ref).max(), c
    fg.close()


def test_resolved_targets_match_oracle(ctx, orc):
    d = lhs_deck.load()
    fg = _fg(ctx, d, 5, _tight())
    th = _sets(d, 5, 2)
    y, st, _ = fg.log_target(th, resolve=True)
    P = lhs_deck.projection(d, d["mesh"].nnodes)
    for c in range(5):
        u, _ = orc.lhs_solution(d["mesh"], d["bc"], d["anchors"], np.exp(th[c, :5]), *d["gp"], nq1d=3)
        yo = lhs_deck.oracle_target(orc, d, th[c], P @ u)
        assert abs(y[c] - yo) <= 1e-8 * abs(yo), c
    fg.close()


def test_resolving_chain_matches_oracle_chain(ctx, orc):
    """Four chains x 40 steps that re-solve every proposal: the same states and
    log targets as the oracle's chain with the same draws."""
    d = lhs_deck.load()
    fg = _fg(ctx, d, 4, _tight())
    th0 = _sets(d, 4, 3)
    samples, acc, st = fg.mcmc(th0, d["propsd"] * 2.0, 40, seed=11, resolve=True)
    P = lhs_deck.projection(d, d["mesh"].nnodes)

    def target(t):
        u, k = orc.lhs_solution(d["mesh"], d["bc"], d["anchors"], np.exp(t[:5]), *d["gp"], nq1d=3)
        if (k <= 0).any():
            return -np.inf
        return lhs_deck.oracle_target(orc, d, t, P @ u)

    for c in range(4):
        ref, a = orc.lhs_chain(th0[c], d["propsd"] * 2.0, 40, 11, c, target)
        assert abs(acc[c] - a) < 1e-12, c
        assert np.abs(samples[c] - ref).max() <= 1e-8 * np.abs(ref).max(), c
    assert 0 < acc.mean() < 1
    fg.close()


def test_reference_chain_statistics(ctx, orc):
    """The reference's own chain (fixed FE solution; init, propStDev and prior of
    the deck; 5000 steps, 30 % burn-in) run as 64 lockstep chains. The
    likelihood does not depend on the anchor values, so their pooled marginal
    means are the log-normal prior's, and the sensor noise's is the
    one-dimensional posterior mean by quadrature. The recorded single-chain
    means (hParMean.dat) lie within the spread of the 64 single-chain
    estimates (the chain is strongly autocorrelated: a few tens of effective
    samples per chain)."""
    d = lhs_deck.load()
    C, steps = 64, 5000
    fg = _fg(ctx, d, C, _tight())
    fg.updateSolution(np.r_[np.log(d["w0"]), np.log(d["hp"]["mismatch.sensorEps"])])
    fg.fix_mean(0)
    th0 = np.repeat(d["init"][None, :], C, axis=0)
    samples, acc, _ = fg.mcmc(th0, d["propsd"], steps, seed=5, resolve=False)
    nat = np.exp(samples[:, steps * 3 // 10:, :6])
    w_mean = nat[:, :, :5].reshape(-1, 5).mean(axis=0)
    w_exact = np.exp(d["pm"][:5] + 0.5 * d["pv"][:5])
    assert np.all(np.abs(w_mean / w_exact - 1.0) < 0.06), (w_mean, w_exact)
    P = lhs_deck.projection(d, d["mesh"].nnodes)
    u, _ = orc.lhs_solution(d["mesh"], d["bc"], d["anchors"], d["w0"], *d["gp"], nq1d=3)
    mean = P @ u
    # log target of log eps alone (the anchor terms factor out)
    ts = np.linspace(np.log(0.006), np.log(0.02), 1500)
    f = np.array([lhs_deck.oracle_target(orc, d, np.r_[d["pm"][:5], t], mean) for t in ts])
    p = np.exp(f - f.max())
    eps_exact = np.sum(np.exp(ts) * p) / np.sum(p)
    assert abs(nat[:, :, 5].mean() / eps_exact - 1.0) < 0.01, (nat[:, :, 5].mean(), eps_exact)
    per_chain = nat.mean(axis=1)                      # 64 x 6 single-chain estimates
    rec = np.array([d["hparmean"][k] for k in lhs_deck.IDS])
    z = (rec - per_chain.mean(axis=0)) / per_chain.std(axis=0)
    assert np.all(np.abs(z) < 4.5), z
    assert 0.5 < acc.mean() < 0.99
    fg.close()


def test_nonpositive_conductivity_is_rejected(ctx, orc):
    """Anchor values whose GP mean dips below zero: updateSolution reports
    SFH_ENOTSPD and the log target of that set is -inf."""
    from statfem_amd import hip
    d = lhs_deck.load()
    fg = _fg(ctx, d, 2, _tight())
    bad = np.log(np.array([10, 1e-3, 10, 1e-3, 10, 0.01]))
    good = d["init"]
    assert orc.lhs_conductivity(d["mesh"], d["anchors"], np.exp(bad[:5]), *d["gp"]).min() < 0
    with pytest.raises(hip.SfhError):
        fg.updateSolution(np.stack([good, bad]))
    y, st, _ = fg.log_target(np.stack([good, bad]), resolve=True)
    assert np.isfinite(y[0]) and y[1] == -np.inf
    fg.close()


def test_two_dimensional_copies_use_multigrid(ctx, orc):
    """A 2D tensor grid (65 x 65 nodes), eight copies: each copy gets its own
    multigrid hierarchy (coarse levels formed on the device every assembly);
    the log targets match the oracle's direct solves."""
    from statfem_amd import mesh as M
    from statfem_amd.lhs import ForwardGaussLHSHip
    m = M.grid2d(64)
    bc = M.dirichlet_all_boundary(m)
    A = np.array([[0.2, 0.2], [0.8, 0.2], [0.5, 0.5], [0.2, 0.8], [0.8, 0.8]])
    rng = np.random.default_rng(4)
    idx = rng.choice(m.nnodes, 24, replace=False)
    cols = [np.array([i], np.int32) for i in idx]
    vals = [np.ones(1) for _ in idx]
    xs = m.coords[idx]
    Y = 0.05 * rng.standard_normal((24, 3)) + 0.02
    fg = ForwardGaussLHSHip(ctx, m, bc, A, (1.0, 0.4, 1e-4), cols, vals, copies=8, nq1d=2, opts=_tight())
    pm = np.r_[np.zeros(5), 0.05]
    pv = np.r_[0.1 * np.ones(5), 0.01]
    fg.set_data(Y, xs, rho=1.0, mis_sigma=0.01, mis_ell=0.2, prior_mu=pm, prior_var=pv)
    th = np.c_[0.3 * rng.standard_normal((8, 5)), np.log(0.05) + 0.1 * rng.standard_normal(8)]
    y, st, stats = fg.log_target(th, resolve=True)
    assert stats["precond"] == 2 and stats["levels"] > 2
    for c in range(8):
        u, _ = orc.lhs_solution(m, bc, A, np.exp(th[c, :5]), 1.0, 0.4, 1e-4, nq1d=2)
        yo = orc.lhs_log_target(th[c], u[idx], Y.T.copy(), xs, 1.0, 0.01, 0.2, pm, pv)
        assert abs(y[c] - yo) <= 1e-8 * abs(yo), c
    fg.close()


def test_laplace_inverse_deck_runs_and_writes_outputs(ctx, tmp_path):
    """The whole application on its own deck: the output files in the
    reference's format (6 significant digits), and the noise mean near the
    recorded one (a loose check: the recorded value is one autocorrelated
    chain; the statistics test above bounds it properly)."""
    from statfem_amd.lhs import laplace_inverse
    from statfem_amd import io
    r = laplace_inverse(ctx, lhs_deck.DECK, out_dir=str(tmp_path), chains=16, steps=3000, seed=1)
    means = io.read_param_map(str(tmp_path / "hParMean.dat"))
    assert list(means) == lhs_deck.IDS
    assert np.allclose([means[k] for k in lhs_deck.IDS], r["means"], rtol=1e-5, atol=0)
    rows = np.loadtxt(str(tmp_path / "logMLsamples.dat"))
    assert rows.shape == (2100, 7)
    assert np.allclose(rows[:, :6], np.exp(r["samples"][0, :, :6]), rtol=1e-5, atol=0)
    assert abs(r["means"][5] / d_rec()["mismatch.sensorEps"] - 1.0) < 0.06


def d_rec():
    return lhs_deck.load()["hparmean"]


def test_log_target_needs_the_data(ctx, orc):
    """Evaluations before sfh_lhs_set_data (or, for resolve = 0, before
    sfh_lhs_fix_mean) are refused, also after an updateSolution."""
    from statfem_amd import hip
    from statfem_amd.lhs import ForwardGaussLHSHip
    d = lhs_deck.load()
    fg = ForwardGaussLHSHip(ctx, d["mesh"], d["bc"], d["anchors"], d["gp"], d["cols"], d["vals"], copies=2, nq1d=3)
    fg.updateSolution(np.repeat(d["init"][None, :], 2, axis=0))
    with pytest.raises(hip.SfhError):
        fg.log_target(np.repeat(d["init"][None, :], 2, axis=0), resolve=True)
    hp = d["hp"]
    fg.set_data(d["Y"], d["xs"], rho=hp["rho"], mis_sigma=hp["mismatch.sigma"], mis_ell=hp["mismatch.length"],
                prior_mu=d["pm"], prior_var=d["pv"])
    with pytest.raises(hip.SfhError):
        fg.log_target(d["init"][None, :], resolve=False)
    y, st, _ = fg.log_target(np.repeat(d["init"][None, :], 2, axis=0), resolve=True)
    assert np.all(st == 0) and np.all(np.isfinite(y))
    fg.close()
