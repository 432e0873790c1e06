"""Chain diagnostics for the samplers: multi-chain effective sample size
(Geyer's initial monotone sequence over the chain-averaged autocorrelation)
and split R-hat. The reference reports only acceptance rates and raw
samples (statX/infer/MCMC.hpp:150-158); these are needed to compare samplers
by effective samples per second."""
import numpy as np


def _autocov(x):
    """Autocovariance of each chain (rows) at every lag, by FFT."""
    m, n = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    nfft = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(xc, nfft, axis=1)
    ac = np.fft.irfft(f * np.conj(f), nfft, axis=1)[:, :n]
    return ac / n


def ess(samples):
    """Effective sample size of one scalar quantity over chains x draws."""
    x = np.asarray(samples, dtype=np.float64)
    if x.ndim == 1:
        x = x[None, :]
    m, n = x.shape
    if n < 4:
        return float(m * n)
    acov = _autocov(x)
    chain_var = acov[:, 0] * n / (n - 1.0)
    w = chain_var.mean()
    b_over_n = x.mean(axis=1).var(ddof=1) if m > 1 else 0.0
    var_plus = w * (n - 1.0) / n + b_over_n
    if var_plus <= 0:
        return float(m * n)
    rho = 1.0 - (w - acov.mean(axis=0)) / var_plus
    rho[0] = 1.0
    # Geyer: sums of adjacent pairs, truncated at the first non-positive pair
    # and made monotone
    pairs = rho[0:n - 1:2][: (n - 1) // 2] + rho[1:n:2][: (n - 1) // 2]
    tau = -1.0
    prev = np.inf
    for p in pairs:
        if p <= 0:
            break
        p = min(p, prev)
        tau += 2.0 * p
        prev = p
    tau = max(tau, 1.0 / np.log10(max(m * n, 10)))
    return float(m * n / tau)


def ess_all(samples):
    """ESS of every parameter of chains x draws x params samples."""
    s = np.asarray(samples)
    return np.array([ess(s[:, :, j]) for j in range(s.shape[2])])


def split_rhat(samples):
    """Split R-hat of one scalar quantity over chains x draws."""
    x = np.asarray(samples, dtype=np.float64)
    if x.ndim == 1:
        x = x[None, :]
    n = x.shape[1] // 2
    x = np.concatenate([x[:, :n], x[:, n:2 * n]], axis=0)
    w = x.var(axis=1, ddof=1).mean()
    b = n * x.mean(axis=1).var(ddof=1)
    return float(np.sqrt(((n - 1.0) / n * w + b / n) / w)) if w > 0 else 1.0


def mcse_mean(samples):
    """Monte-Carlo standard error of the posterior mean of each parameter."""
    s = np.asarray(samples)
    return s.reshape(-1, s.shape[2]).std(axis=0) / np.sqrt(ess_all(s))
