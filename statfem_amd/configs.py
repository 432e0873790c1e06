"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

The reference ships no samplers for these fields; the choices here are the
documented stand-ins: Gaussian-process samples on tensor grids are drawn
through the Kronecker factorisation K = K_y (x) K_x with eigen-factors
Q sqrt(max(Lambda, 0)) of each 1D squared-exponential matrix.
"""
import numpy as np

from . import mesh as M


def se_1d(x, ell):
    d = x[:, None] - x[None, :]
    return np.exp(-d * d / (2.0 * ell * ell))


def _sqrt_factor(K):
    lam, Q = np.linalg.eigh(K)
    return Q * np.sqrt(np.clip(lam, 0.0, None))[None, :]


def gp_sample_grid(axes, sigma, ell, rng):
    """Sample of a zero-mean SE GP on the tensor grid of `axes` (x fastest)."""
    L = [_sqrt_factor(se_1d(a, ell)) for a in axes]
    shape = [len(a) for a in axes][::-1]          # (..., ny, nx)
    Z = rng.standard_normal(shape)
    if len(axes) == 1:
        g = L[0] @ Z
    elif len(axes) == 2:
        g = L[1] @ Z @ L[0].T
    else:
        g = np.einsum("ia,jb,kc,abc->ijk", L[2], L[1], L[0], Z)
    return sigma * g.ravel()


def gp_sample_points(xs, sigma, ell, rng, nugget=1e-12):
    d = xs[:, None, :] - xs[None, :, :]
    K = sigma * sigma * np.exp(-(d ** 2).sum(-1) / (2 * ell * ell))
    L = np.linalg.cholesky(K + nugget * sigma * sigma * np.eye(len(xs)))
    return L @ rng.standard_normal(len(xs))


def c1_inputs(seed=0, n=64, ns=32, no=10):
    """Plain statFEM Poisson: unit square, n x n grid split into P1 triangles,
    homogeneous Dirichlet, f = 1 + GP(0.1, 0.2) on nodes, 32 interior sensors,
    discrepancy GP(0.05, 0.3), noise sigma_e = 1e-3, rho = 1, `no` snapshots."""
    rng = np.random.default_rng(seed)
    m = M.grid2d(n, kind=M.TRI3)
    bc = M.dirichlet_all_boundary(m)
    g = gp_sample_grid(m.grid, 0.1, 0.2, rng)
    xs = rng.uniform(0.0, 1.0, size=(ns, 2)) * (1 - 2.0 / n) + 1.0 / n
    cols, vals = M.sensor_rows(m, xs)
    d = np.stack([gp_sample_points(xs, 0.05, 0.3, rng) for _ in range(no)])
    e = rng.standard_normal((no, ns)) * 1e-3
    return dict(mesh=m, bc=bc, f_true=1.0 + g, xs=xs, cols=cols, vals=vals, d=d, e=e, rho=1.0,
                sigma_f=0.1, ell_f=0.2, sigma_d=0.05, ell_d=0.3, sigma_e=1e-3, nq1d=1, reg=1e-3)


def kappa_field(n, seed, ell=0.1, sigma=0.5):
    """Element diffusivity exp(g), g ~ GP(SE ell, sigma) on the n x n centroids."""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n
    return np.exp(gp_sample_grid((c, c), sigma, ell, rng))


def c2_inputs(seed=0, n=512, ns=256):
    """Poisson with many right-hand sides: n x n Q1, kappa = exp(GP(0.1, 0.5)) per
    element, forcing covariance SE(0.1, 0.05), `ns` uniform sensors."""
    m = M.grid2d(n, kind=M.QUAD4)
    bc = M.dirichlet_all_boundary(m)
    kappa = kappa_field(n, seed)
    rng = np.random.default_rng(seed + 1)
    xs = rng.uniform(0.0, 1.0, size=(ns, 2))
    cols, vals = M.sensor_rows(m, xs)
    return dict(mesh=m, bc=bc, kappa=kappa, xs=xs, cols=cols, vals=vals, sigma_f=0.1, ell_f=0.05,
                nq1d=2, reg=1e-3)


def c5_inputs(seed=0, n=256, ns=200, no=8):
    """Hyperparameter MCMC problem: Poisson on n x n Q1 (kappa = 1), homogeneous
    Dirichlet, mean forcing 1 with SE forcing covariance (0.1, 0.2), `ns`
    interior sensors, `no` observation snapshots drawn from the statFEM data
    model at the true hyperparameters (rho, sigma_d, ell_d) = (1.2, 0.05, 0.25)
    with noise sigma_e = 1e-3. Log-normal priors: medians (1, 0.1, 0.3),
    log-variance 0.5 each."""
    m = M.grid2d(n, kind=M.QUAD4)
    bc = M.dirichlet_all_boundary(m)
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0.0, 1.0, size=(ns, 2)) * (1 - 2.0 / n) + 1.0 / n
    cols, vals = M.sensor_rows(m, xs)
    return dict(mesh=m, bc=bc, xs=xs, cols=cols, vals=vals, no=no, sigma_f=0.1, ell_f=0.2, nq1d=2, reg=1e-3,
                truth=np.array([1.2, 0.05, 0.25]), sigma_e=1e-3, data_seed=seed + 1000,
                prior_mu=np.log([1.0, 0.1, 0.3]), prior_var=np.array([0.5, 0.5, 0.5]))


def c3_inputs(seed=0, n=64, ns=128):
    """3D linear elasticity: unit cube of n^3 HEX8 cells, E = 1, nu = 0.3, the
    z = 0 face clamped. Vertical body force f_z = -1 + g with g a Matern field
    (SPDE alpha = 2, so nu_m = 1/2 in 3D, ell = 0.15, kappa = sqrt(8 nu_m)/ell,
    sigma_g = 0.1). `ns` sensors of the vertical displacement at seeded points
    of the top face, noise sigma_e = 1e-4."""
    m = M.grid3d(n)
    bottom = M.fuzzy_boundary(m.coords, [2], sides=("lo",))
    bc = np.sort(np.concatenate([3 * bottom, 3 * bottom + 1, 3 * bottom + 2])).astype(np.int32)
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, 1.0, size=(ns, 2))
    pts = np.concatenate([xy, np.ones((ns, 1))], axis=1)
    cols, vals = M.sensor_rows(m, pts)
    nu_m = 2.0 - 1.5
    ell = 0.15
    return dict(mesh=m, bc=bc, bottom=bottom, xs=pts, cols=3 * cols + 2, vals=vals, E=1.0, nu=0.3,
                fbar=-1.0, sigma_g=0.1, ell_g=ell, nu_m=nu_m, kappa=np.sqrt(8.0 * nu_m) / ell, sigma_e=1e-4,
                nq1d=2, noise_seed=seed + 77)


def c4_inputs(seed=0, nprob=64, nmin=32, nmax=256, smin=16, smax=128):
    """Batched uneven statFEM: `nprob` independent Poisson problems on the unit
    square with n x n Q1 cells, n ~ U{nmin..nmax}; forcing mean 1 with SE
    covariance (sigma_f = 0.1, ell_f ~ U(0.05, 0.3)); n_s ~ U{smin..smax}
    sensors at uniform interior points; homogeneous Dirichlet boundary."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(nprob):
        n = int(rng.integers(nmin, nmax + 1))
        ell = float(rng.uniform(0.05, 0.3))
        ns = int(rng.integers(smin, smax + 1))
        m = M.grid2d(n, kind=M.QUAD4)
        xs = rng.uniform(0.0, 1.0, size=(ns, 2)) * (1 - 2.0 / n) + 1.0 / n
        cols, vals = M.sensor_rows(m, xs)
        probs.append(dict(n=n, mesh=m, bc=M.dirichlet_all_boundary(m), xs=xs, cols=cols, vals=vals,
                          sigma_f=0.1, ell_f=ell, reg=1e-3, sigma_e=1e-3))
    return probs


def lhs_inputs(seed=0, n=128, ns=64, no=5):
    """Conductivity (LHS) inverse problem in the laplaceInverse form on a 2D grid:
    Poisson on n x n Q1 with unit body force and homogeneous Dirichlet on the
    whole boundary; the conductivity of each element is the GP mean
    (sigma 1, length 0.4, nugget 5e-5) at its centroid of 5 anchor values at
    the corners of the inner square [0.2, 0.8]^2 and its centre; true anchor
    values (0.7, 1.0, 0.7, 0.4, 0.7), `ns` interior sensors, `no` readings
    with noise sigma_e = 0.01 drawn from the data model (the readings are made
    by the caller from a forward solve at the true values). Log-normal priors
    on the anchors (medians 0.8, 1.1, 0.8, 0.4, 0.8, log-variance 0.05), normal
    prior on sigma_e (0.0075, 0.002), proposal SDs 0.02 and 0.01, as the
    reference deck."""
    m = M.grid2d(n, kind=M.QUAD4)
    bc = M.dirichlet_all_boundary(m)
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0.0, 1.0, size=(ns, 2)) * (1 - 2.0 / n) + 1.0 / n
    cols, vals = M.sensor_rows(m, xs)
    anchors = np.array([[0.2, 0.2], [0.8, 0.2], [0.5, 0.5], [0.2, 0.8], [0.8, 0.8]])
    return dict(mesh=m, bc=bc, xs=xs, cols=cols, vals=vals, no=no, anchors=anchors, gp=(1.0, 0.4, 5e-5),
                w_true=np.array([0.7, 1.0, 0.7, 0.4, 0.7]), sigma_e=0.01, noise_seed=seed + 77,
                prior_mu=np.r_[np.log([0.8, 1.1, 0.8, 0.4, 0.8]), 0.0075], prior_var=np.r_[0.05 * np.ones(5), 0.002],
                init=np.log([0.8, 1.1, 0.8, 0.4, 0.8, 0.0075]), propsd=np.r_[0.02 * np.ones(5), 0.01])
