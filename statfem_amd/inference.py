"""Hyperparameter inference drivers over the GPU likelihood.

`tuned_proposal` follows the reference's random-walk Metropolis sampler
(statX/infer/MCMC.hpp:123-158) with a pilot run that scales the step sizes
to the posterior. `surrogate_posterior` adds the learned screen: a training
design around the pilot posterior, exact log-likelihoods for it from the GPU
likelihood, GPU training of the MLP, then delayed-acceptance sampling whose
exact second stage keeps the exact posterior."""
import time

import numpy as np

from . import diagnostics as D
from . import hip


def tuned_proposal(lik, theta0, steps=2000, seed=0, init_sd=0.02, **prior):
    """Pilot chains from theta0 (chains x 3, log space); returns the
    random-walk step sizes 2.38 / sqrt(3) x posterior sd, the pilot's last
    states and its wall time."""
    t0 = time.perf_counter()
    s, _ = lik.mcmc(theta0, [init_sd] * 3, steps // 2, seed=seed, **prior)
    sd = s[:, steps // 4:, :3].reshape(-1, 3).std(0)
    s, _ = lik.mcmc(s[:, -1, :3].copy(), 1.4 * sd, steps - steps // 2, seed=seed + 1, **prior)
    post = s[:, (steps - steps // 2) // 2:, :3].reshape(-1, 3)
    return 1.374 * post.std(0), s[:, -1, :3].copy(), post.mean(0), post.std(0), time.perf_counter() - t0


def surrogate_posterior(lik, theta0, propsd, center, scale, steps, sub, seed=0, n_train=4096, train_steps=2000,
                        lr=3e-3, width=2.5, net_seed=0, **prior):
    """Surrogate-screened sampling. The training design is `n_train` points
    center + width * scale * N(0, I) in log space, scored by the exact GPU
    likelihood. Returns samples (chains x steps x 4), acceptance rates and
    the wall time of each phase."""
    ctx = lik.ctx
    t = {}
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    design = center + width * scale * rng.standard_normal((n_train, 3))
    ll, st = lik.giveLogLikelihood(np.exp(design))
    ok = st == 0
    t["design"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    net = hip.SurrogateMLP(ctx, seed=net_seed)
    loss = net.train(design[ok], ll[ok], steps=train_steps, lr=lr)
    t["train"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    samples, acc, sacc = hip.mcmc_da(lik, net, theta0, propsd, steps, sub, seed=seed + 17, **prior)
    t["sample"] = time.perf_counter() - t0
    return dict(samples=samples, acc=acc, sub_acc=sacc, times=t, train_loss=loss, net=net, n_train=int(ok.sum()))


def moment_errors(a, b):
    """Relative error of the natural-scale posterior mean and sd of chains a
    against chains b (chains x draws x >=3, log space)."""
    xa = np.exp(a[:, :, :3].reshape(-1, 3))
    xb = np.exp(b[:, :, :3].reshape(-1, 3))
    return np.abs(xa.mean(0) / xb.mean(0) - 1), np.abs(xa.std(0) / xb.std(0) - 1)


def ess_per_second(samples, seconds, burn=0.1):
    s = samples[:, int(burn * samples.shape[1]):, :3]
    e = D.ess_all(s)
    return e, e.min() / seconds
