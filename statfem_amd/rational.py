"""Best rational approximation of x^gamma for the fractional SPDE exponent.

PrecisionMatrixRational::brasil_ (statX/inferQ/PrecisionMatrixRational.ipp:300-344)
runs tools/baryrat/baryrat-1.4.0/roots.py, which calls baryrat 1.4.0's
`brasil(f, [delta, 1], [m, m])` on f(x) = x^gamma and converts the result to
factor form (interfaceBrasil.py:27-58):

    r(x) = a_m prod_i (x - c_i) / (b_m prod_j (x - d_j)),
    top = [a_m, c_0 .. c_{m-1}],  bot = [b_m, d_0 .. d_{m-1}],

with delta = 1e-15 for gamma > 0 and 10^(-(5 + m) / 2) for gamma < 0, and the
leading coefficients read off at x = 1/2. baryrat is a third-party package
that is not part of the reference; this module restates the published
algorithm it implements (C. Hofreither, "An algorithm for best rational
approximation based on barycentric rational interpolation", Numerical
Algorithms 88 (2021) 365-388):

- a type (m, m) barycentric interpolant through 2m + 1 nodes: the even nodes
  are support points, the odd ones fix the weights as the null vector of the
  Loewner matrix of divided differences;
- the 2m + 2 intervals between the end points and the nodes each hold one
  local maximum of |f - r|; the best approximation equioscillates, so all of
  them are equal (Chebyshev's alternation theorem);
- each step rescales the interval lengths by a power of (mean error / local
  error), shrinking intervals whose error is above the mean, until the local
  maxima agree to `tol` relative.

The best approximation is unique, so any implementation converged to a
tolerance agrees with baryrat's to that tolerance; baryrat's default stops at
a relative spread of 1e-4 and this one goes to 1e-8 by default. Only the
ratio a_m / b_m enters the precision (Q_v scales with b_m^2, P_r with a_m),
so the normalisation of the weights does not matter."""
import numpy as np
import scipy.linalg as sla


def _delta(exponent, degree):
    """Left end of the interval (interfaceBrasil.py:27-30)."""
    return 1.0e-15 if exponent > 0.0 else 10.0 ** (-(5.0 + degree) / 2.0)


class _Bary:
    """r(x) = sum w_i f_i / (x - s_i) / sum w_i / (x - s_i) on support points s."""

    def __init__(self, s, fs, w):
        self.s, self.fs, self.w = s, fs, w

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        d = x[..., None] - self.s
        exact = d == 0.0
        d = np.where(exact, 1.0, d)
        c = self.w / d
        r = (c * self.fs).sum(-1) / c.sum(-1)
        hit = exact.any(-1)
        if np.any(hit):
            r = np.where(hit, (exact * self.fs).sum(-1), r)
        return r

    def _pencil_roots(self, vals):
        # finite eigenvalues of [[0, (w v)^T], [1, diag(s)]] - lambda diag(0, 1, ..., 1)
        n = len(self.s)
        E = np.zeros((n + 1, n + 1))
        E[0, 1:] = self.w * vals
        E[1:, 0] = 1.0
        E[1:, 1:] = np.diag(self.s)
        B = np.eye(n + 1)
        B[0, 0] = 0.0
        ev = sla.eigvals(E, B)
        ev = ev[np.isfinite(ev)]
        ev = ev[np.argsort(np.abs(ev))][: n - 1]
        return np.real_if_close(ev, tol=1e6)

    def _polish(self, roots, vals, steps=8):
        """Newton steps on g(x) = sum_i w_i v_i / (x - s_i), whose roots are the
        pencil's (simple and real here)."""
        x = np.real(np.asarray(roots, np.complex128)).copy()
        for _ in range(steps):
            d = x[:, None] - self.s
            if np.any(d == 0.0):
                break
            g = (self.w * vals / d).sum(-1)
            dg = -(self.w * vals / d ** 2).sum(-1)
            ok = dg != 0.0
            x = np.where(ok, x - np.where(ok, g / np.where(ok, dg, 1.0), 0.0), x)
        return x

    def zeros(self):
        return self._polish(self._pencil_roots(self.fs), self.fs)

    def poles(self):
        one = np.ones_like(self.fs)
        return self._polish(self._pencil_roots(one), one)

    def numerator(self, x):
        """sum_i w_i f_i prod_{k != i} (x - s_k)"""
        return sum(self.w[i] * self.fs[i] * np.prod(np.delete(x - self.s, i)) for i in range(len(self.s)))

    def denominator(self, x):
        return sum(self.w[i] * np.prod(np.delete(x - self.s, i)) for i in range(len(self.s)))


def _interpolant(f, nodes):
    s, t = nodes[0::2], nodes[1::2]
    fs, ft = f(s), f(t)
    L = (ft[:, None] - fs[None, :]) / (t[:, None] - s[None, :])
    w = np.linalg.svd(L)[2][-1]
    return _Bary(s, fs, w)


def _local_maxima(err, edges, nsamp=64, iters=40):
    """Largest |err| on each interval [edges[i], edges[i+1]] (end points
    included only at the two ends): sampled, geometrically where the interval
    spans more than a factor 16, then refined by golden-section search around
    the best sample, all intervals at once."""
    n = len(edges) - 1
    t = np.linspace(0.0, 1.0, nsamp)
    lo, hi = edges[:-1, None], edges[1:, None]
    geo = (lo > 0.0) & (hi > 16.0 * lo)
    x = np.where(geo, np.exp(np.log(np.where(geo, lo, 1.0)) * (1.0 - t) + np.log(np.where(geo, hi, 1.0)) * t),
                 lo + (hi - lo) * t)
    e = np.abs(err(x))
    e[1:, 0] = -1.0        # interior nodes: the interpolant is exact there
    e[:-1, -1] = -1.0
    k = np.argmax(e, axis=1)
    rows = np.arange(n)
    best, best_x = e[rows, k], x[rows, k]
    a = x[rows, np.maximum(k - 1, 0)]
    b = x[rows, np.minimum(k + 1, nsamp - 1)]
    g = (np.sqrt(5.0) - 1.0) / 2.0
    for _ in range(iters):
        c = b - g * (b - a)
        d = a + g * (b - a)
        ec, ed = np.abs(err(c)), np.abs(err(d))
        left = ec >= ed
        b = np.where(left, d, b)
        a = np.where(left, a, c)
        cand = np.where(left, ec, ed)
        cx = np.where(left, c, d)
        up = cand > best
        best = np.where(up, cand, best)
        best_x = np.where(up, cx, best_x)
    return best, best_x


def brasil(f, interval, degree, tol=1e-8, maxiter=2000, step=0.5, max_factor=2.0):
    """Best (degree, degree) rational approximation of f on `interval`.
    Returns (r, info) with info = dict(iters, spread, errors, extrema)."""
    a, b = float(interval[0]), float(interval[1])
    n = 2 * degree + 1
    # start from Chebyshev nodes graded towards the left end: geometrically
    # between a and b when a > 0 (the singular end of x^gamma), else cubically
    k = np.arange(n)
    u = (1.0 - np.cos(np.pi * (k + 0.5) / n)) / 2.0
    nodes = a * (b / a) ** u if a > 0.0 else a + (b - a) * u ** 3
    L = np.diff(np.r_[a, nodes, b])
    best = None
    spread = np.inf
    stall = 0
    for it in range(maxiter):
        nodes = a + np.cumsum(L)[:-1]
        r = _interpolant(f, nodes)
        err = lambda x, r=r: f(x) - r(x)
        E, X = _local_maxima(err, np.r_[a, nodes, b])
        spread = (E.max() - E.min()) / E.max()
        if best is None or spread < best[1]:
            best = (r, spread, E.copy(), X.copy(), it)
            stall = 0
        else:
            stall += 1
            if stall >= 50:   # no progress: smaller steps
                step *= 0.5
                stall = 0
        if spread < tol:
            break
        mean = np.exp(np.log(E).mean())
        fac = np.clip((mean / E) ** step, 1.0 / max_factor, max_factor)
        L = L * fac
        L *= (b - a) / L.sum()
    r, spread, E, X, it = best
    return r, {"iters": it, "spread": spread, "errors": E, "extrema": X}


def brasil_factors(exponent, degree=4, tol=1e-8, maxiter=2000):
    """(top, bot) of interfaceBrasil.rational (interfaceBrasil.py:39-58) for
    f(x) = x^exponent on [delta, 1]: leading coefficient then the degree roots
    of the numerator, and of the denominator."""
    if not (0.0 < abs(exponent) < 1.0):
        raise ValueError("the exponent must lie in (-1, 0) or (0, 1)")
    f = lambda x: np.asarray(x, np.float64) ** exponent
    r, info = brasil(f, [_delta(exponent, degree), 1.0], degree, tol=tol, maxiter=maxiter)
    if not info["spread"] < 1e-4:   # baryrat's own default tolerance
        raise RuntimeError(f"best rational approximation of x^{exponent} (degree {degree}) did not converge: "
                           f"relative spread of the error maxima {info['spread']:.2e}")
    c = np.sort(r.zeros())
    d = np.sort(r.poles())
    x0 = 0.5
    a = r.numerator(x0) / np.prod(x0 - c)
    bm = r.denominator(x0) / np.prod(x0 - d)
    return np.r_[a, c], np.r_[bm, d], info


def factor_form(top, bot, x):
    """a_m prod (x - c_i) / (b_m prod (x - d_j)) at x."""
    x = np.asarray(x, np.float64)[..., None]
    return top[0] * np.prod(x - top[1:], -1) / (bot[0] * np.prod(x - bot[1:], -1))
