"""Sparse-precision (SPDE) Gaussian-process regression on the GPU:
GaussianProcessQ (statX/inferQ/GaussianProcessQ.hpp:65-159,
GaussianProcessQ.ipp:37-186) with the precision that DriverMaxLikelihood
assembles (statX/inferQ/DriverMaxLikelihood.hpp:52-313). A fractional beta
uses the rational factors of PrecisionMatrixRational (statX/inferQ/
PrecisionMatrixRational.ipp:24-161); they come from `rational.brasil_factors`
for the exponent beta - floor(beta) at the reference's degree 4 (its
brasil_ call, :300-344), or from `set_rational`.

The likelihood and posterior go through the sensor-space matrix
G = P Q^-1 P^T, formed from multi-right-hand-side solves with the Matern
operator (sfh_gpq_* in include/statfem_hip.h), instead of the reference's two
sparse Cholesky factorisations per evaluation."""
import ctypes as C

import numpy as np

from . import hip
from . import rational


def _rows(cols, vals):
    ns = len(cols)
    pn = max(len(c) for c in cols)
    pc = np.zeros((ns, pn), np.int32)
    pv = np.zeros((ns, pn), np.float64)
    for i, (c, v) in enumerate(zip(cols, vals)):
        pc[i, :len(c)] = c
        pv[i, :len(c)] = v
        if len(c) < pn:
            pc[i, len(c):] = c[0] if len(c) else 0
    return pc, pv, pn


class GaussianProcessQHip:
    """obsValues Y (ns x no, as the reference stores them) and the projection
    P (rows of (DOF columns, values)) on `system`'s scalar mesh."""

    def __init__(self, system, cols, vals, Y, nq1d=2, degree=4):
        self.system = system
        self.ctx = system.ctx
        self.dim = system.mesh.dim
        self.degree = degree
        self.rationals = {}
        self.ns = len(cols)
        Y = np.asarray(Y, np.float64).reshape(self.ns, -1)
        self.no = Y.shape[1]
        pc, pv, pn = _rows(cols, vals)
        Yr = np.ascontiguousarray(Y.T)
        h = C.c_void_p()
        hip._check(hip.lib().sfh_gpq_create(system.h, nq1d, self.ns, pn, hip._p(pc), hip._p(pv), self.no,
                                            hip._p(Yr), C.byref(h)))
        self.h = h
        self.ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            hip.lib().sfh_gpq_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def exponent(self, nu):
        """beta - alpha with beta = (nu + d/2) / 2, alpha = max(1, floor beta)
        (PrecisionMatrixRational.ipp:64-67); 0 for integer beta."""
        beta = (nu + self.dim / 2.0) / 2.0
        e = beta - max(1.0, np.floor(beta))
        return 0.0 if abs(e) < 1e-12 else e

    def set_rational(self, exponent, top, bot):
        """The rational factors for one exponent: top = (a, c_0 ..), bot = (b, d_0 ..)."""
        top = hip._h(np.asarray(top, np.float64))
        bot = hip._h(np.asarray(bot, np.float64))
        if len(top) != len(bot) or len(top) < 2:
            raise ValueError("top and bot need the same degree + 1 entries")
        hip._check(hip.lib().sfh_gpq_set_rational(self.h, float(exponent), len(top) - 1, hip._p(top), hip._p(bot)))
        self.rationals[float(exponent)] = (top, bot)

    def _rationals_for(self, nus):
        for nu in nus:
            e = self.exponent(float(nu))
            if e != 0.0 and not any(abs(e - k) < 1e-12 for k in self.rationals):
                top, bot, _ = rational.brasil_factors(e, self.degree)
                self.set_rational(e, top, bot)

    def giveLogLikelihood(self, hps, opts=None):
        """hps: M x 4 sets (sigma, kappa, nu, eps). Returns (logL, status, stats)."""
        t = np.ascontiguousarray(np.asarray(hps, np.float64).reshape(-1, 4))
        self._rationals_for(np.unique(t[:, 2]))
        out = np.zeros(t.shape[0])
        st = np.zeros(t.shape[0], np.int32)
        ss = hip.SolveStats()
        o = opts or hip.default_opts()
        hip._check(hip.lib().sfh_gpq_loglik(self.h, t.shape[0], hip._p(t), C.byref(o), hip._p(out), hip._p(st),
                                            C.byref(ss)))
        return out, st, ss.as_dict()

    def _predict(self, hp, mean, cov, opts):
        t = hip._h(np.asarray(hp, np.float64).reshape(4))
        self._rationals_for([t[2]])
        n = self.system.ndof
        md = self.ctx.empty((n,)) if mean else None
        cd = self.ctx.empty((n, n)) if cov else None
        ll = C.c_double(0.0)
        ss = hip.SolveStats()
        o = opts or hip.default_opts()
        hip._check(hip.lib().sfh_gpq_predict(self.h, hip._p(t), C.byref(o), None if md is None else md.ptr,
                                             None if cd is None else cd.ptr, C.byref(ll), C.byref(ss)))
        return md, cd, ll.value, ss.as_dict()

    def predictMean(self, hp, opts=None):
        md, _, _, _ = self._predict(hp, True, False, opts)
        return md.download()

    def predictCovariance(self, hp, opts=None):
        _, cd, _, _ = self._predict(hp, False, True, opts)
        return cd.download()

    def predictVariance(self, hp, opts=None):
        return np.diag(self.predictCovariance(hp, opts)).copy()


class DriverMaxLikelihoodHip:
    """DriverMaxLikelihood (statX/inferQ/DriverMaxLikelihood.hpp:52-313) over a
    hyperparameter map with the reference's keys: gaussianProcess.sigma, .nu,
    .kappa (or .length, kappa = sqrt(2 nu) / length, :288-312) and .eps."""

    def __init__(self, gp):
        self.gp = gp
        self.hps = {}
        self.names = []

    def setAllHps(self, hps):
        self.hps = dict(hps)

    def giveAllHps(self):
        return dict(self.hps)

    def addOptimHps(self, name):
        self.names.append(name)

    def updateHps(self, params):
        for k, v in zip(self.names, params):
            self.hps[k] = float(v)

    def _set(self, hps=None):
        h = self.hps if hps is None else hps
        nu = h["gaussianProcess.nu"]
        kap = h["gaussianProcess.kappa"] if "gaussianProcess.kappa" in h else np.sqrt(2.0 * nu) / h["gaussianProcess.length"]
        return [h["gaussianProcess.sigma"], kap, nu, h["gaussianProcess.eps"]]

    def computeLogLikelihood(self, opts=None):
        ll, st, _ = self.gp.giveLogLikelihood([self._set()], opts)
        if st[0] != 0:
            raise hip.SfhError(int(st[0]), "log-likelihood failed")
        return float(ll[0])

    def computeLogLikelihoods(self, param_sets, opts=None):
        """Many candidate parameter vectors (for the names added with
        addOptimHps) in one call: sets sharing (sigma, kappa, nu) share their
        solves. Returns (logL, status)."""
        rows = []
        for p in param_sets:
            h = dict(self.hps)
            for k, v in zip(self.names, p):
                h[k] = float(v)
            rows.append(self._set(h))
        ll, st, _ = self.gp.giveLogLikelihood(rows, opts)
        return ll, st

    def predictMean(self, opts=None):
        return self.gp.predictMean(self._set(), opts)

    def predictCovariance(self, opts=None):
        return self.gp.predictCovariance(self._set(), opts)
