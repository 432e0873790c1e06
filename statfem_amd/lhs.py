"""Conductivity (LHS) uncertainty on the GPU: ForwardGaussLHS with the solve
function of laplaceInverse, its likelihood, and the MCMC of its
hyperParamLearning.

Reference: statX/infer/ForwardGaussLHS.hpp:41-105;
del2/apps/reference/laplaceInverse/laplaceInverse.cpp:62-170 (set-up),
inference.hpp:56-164 (solve function, likelihood), hyperParamLearning.hpp:84-266
(MCMC over the anchor conductivities and the sensor noise).

The reference evaluates one conductivity field per likelihood call. Here a
ForwardGaussLHSHip holds `copies` copies of the mesh as one block-diagonal
system, so a batch of hyperparameter sets (one per MCMC chain) is
interpolated, assembled, solved and projected in one pass of batched kernels
(sfh_lhs_* in include/statfem_hip.h). The reference's laplaceInverse computes
the FE solution once, before its chain (inference.hpp:139), and its likelihood
reuses it at every step; `resolve=False` keeps that behaviour, `resolve=True`
re-solves every proposal, which is what ForwardGaussLHS::updateSolution is for.
"""
import ctypes as C
import os

import numpy as np

from . import hip
from . import io
from . import mesh as M

FREE_IDS = ["gaussianProcess.w0", "gaussianProcess.w1", "gaussianProcess.w2", "gaussianProcess.w3",
            "gaussianProcess.w4", "mismatch.sensorEps"]


class ForwardGaussLHSHip:
    """ForwardGaussLHS (statX/infer/ForwardGaussLHS.hpp:41-105) on `copies`
    copies of `mesh`.

    mesh: scalar mesh; bc: its constrained DOFs (homogeneous, fem.hpp:31-43);
    anchors: na x dim GP training coordinates, gp = (sigma, length, eps) of the
    conductivity GaussianProcess; cols/vals: the projection P row by row;
    fconst: the body force (fem.hpp:19-27). A 2D tensor-grid mesh (mesh.grid)
    solves with the multigrid preconditioner, one hierarchy per copy."""

    def __init__(self, ctx, mesh, bc, anchors, gp, cols, vals, copies=1, nq1d=2, fconst=1.0, opts=None):
        self.ctx = ctx
        self.mesh = mesh
        self.copies = copies
        self.nq1d = nq1d
        self.anchors = np.ascontiguousarray(np.asarray(anchors, np.float64).reshape(-1, mesh.dim))
        self.na = self.anchors.shape[0]
        self.np = self.na + 1
        self.ns = len(cols)
        self.opts = opts or hip.default_opts()
        um, off = M.union([mesh] * copies)
        self.seg = np.ascontiguousarray(off, dtype=np.int32)
        s = hip.SystemSolveHip(ctx, um)
        if mesh.grid is not None and mesh.dim == 2:
            nx = np.full(copies, len(mesh.grid[0]), np.int32)
            ny = np.full(copies, len(mesh.grid[1]), np.int32)
            hip._check(hip.lib().sfh_system_set_segment_grids(s.h, copies, hip._p(self.seg), hip._p(nx), hip._p(ny)))
        ubc = np.concatenate([np.asarray(bc, np.int32) + off[c] for c in range(copies)]).astype(np.int32)
        s.assemble(hip.POISSON, 1.0, nq1d=nq1d)
        s.assemble_rhs(nq1d=nq1d, fconst=fconst)
        s.applyConstraints(ubc)
        self.sys = s
        pn = max(len(c) for c in cols)
        pc = np.zeros((self.ns, pn), np.int32)
        pv = np.zeros((self.ns, pn), np.float64)
        for i, (c, v) in enumerate(zip(cols, vals)):
            pc[i, :len(c)] = c
            pv[i, :len(c)] = v
            if len(c) < pn:
                pc[i, len(c):] = c[0] if len(c) else 0
        h = C.c_void_p()
        hip._check(hip.lib().sfh_lhs_create(s.h, copies, hip._p(self.seg), mesh.nel, self.na, hip._p(self.anchors),
                                            float(gp[0]), float(gp[1]), float(gp[2]), nq1d, len(ubc),
                                            hip._p(ubc), self.ns, pn, hip._p(pc), hip._p(pv), C.byref(h)))
        self.h = h
        ctx.adopt(self)
        self._mean = None

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            hip.lib().sfh_lhs_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def _theta(self, theta, rows):
        t = np.ascontiguousarray(np.asarray(theta, np.float64).reshape(-1, self.np))
        if t.shape[0] == 1 and rows > 1:
            t = np.ascontiguousarray(np.repeat(t, rows, axis=0))
        if t.shape[0] != rows:
            raise ValueError(f"{t.shape[0]} hyperparameter sets for {rows}")
        return t

    def updateSolution(self, theta, opts=None):
        """Projected means P A(w)^-1 f of every copy (ForwardGaussLHS.hpp:83-92);
        theta: copies x (na + 1) in log space, or one set for all copies."""
        t = self._theta(theta, self.copies)
        mean = np.zeros((self.copies, self.ns))
        st = hip.SolveStats()
        o = opts or self.opts
        hip._check(hip.lib().sfh_lhs_update(self.h, hip._p(t), C.byref(o), hip._p(mean), C.byref(st)))
        self._mean = mean
        return mean, st.as_dict()

    def giveProjectedMeanU(self, copy=0):
        if self._mean is None:
            raise RuntimeError("updateSolution first")
        return self._mean[copy].copy()

    def giveProjectedCovarU(self):
        """Zero, as in the reference (ForwardGaussLHS.hpp:71-79)."""
        return np.zeros((self.ns, self.ns))

    def set_data(self, Y, xs, rho=1.0, mis_sigma=0.0, mis_ell=1.0, prior_mu=None, prior_var=None):
        """Readings Y (ns x no, the reference's sensorValues layout), sensor
        coordinates xs, LikelihoodKOH's rho and MismatchCovariance parameters,
        and the enrichWithPrior prior (means and variances, na + 1 each)."""
        Y = np.asarray(Y, np.float64).reshape(self.ns, -1)
        Yr = np.ascontiguousarray(Y.T)
        xs = np.ascontiguousarray(np.asarray(xs, np.float64).reshape(self.ns, -1))
        prior = prior_mu is not None
        pm = hip._h(prior_mu) if prior else None
        pv = hip._h(prior_var) if prior else None
        hip._check(hip.lib().sfh_lhs_set_data(self.h, Yr.shape[0], hip._p(Yr), xs.shape[1], hip._p(xs), float(rho),
                                              float(mis_sigma), float(mis_ell), int(prior), hip._p(pm), hip._p(pv)))

    def fix_mean(self, copy=0):
        hip._check(hip.lib().sfh_lhs_fix_mean(self.h, copy))

    def log_target(self, theta, resolve, opts=None):
        """posteriorGivenData (hyperParamLearning.hpp:172-190) of every set in theta."""
        t = self._theta(theta, self.copies if resolve else np.asarray(theta).reshape(-1, self.np).shape[0])
        out = np.zeros(t.shape[0])
        stat = np.zeros(t.shape[0], np.int32)
        st = hip.SolveStats()
        o = opts or self.opts
        hip._check(hip.lib().sfh_lhs_log_target(self.h, t.shape[0], hip._p(t), int(resolve), C.byref(o), hip._p(out),
                                                hip._p(stat), C.byref(st)))
        return out, stat, st.as_dict()

    def mcmc(self, theta0, propsd, steps, seed=0, resolve=True, opts=None, keep_samples=True):
        """MCMC<na + 1>::sample (statX/infer/MCMC.hpp:123-158) for one chain per copy.
        Returns (samples copies x steps x (na + 2) or None, acceptance, solve stats)."""
        t0 = self._theta(theta0, self.copies)
        sd = hip._h(np.broadcast_to(np.asarray(propsd, np.float64), (self.np,)))
        samples = np.zeros((self.copies, steps, self.np + 1)) if keep_samples else None
        acc = np.zeros(self.copies)
        st = hip.SolveStats()
        o = opts or self.opts
        hip._check(hip.lib().sfh_lhs_mcmc(self.h, steps, hip._p(t0), hip._p(sd), C.c_uint64(seed), int(resolve),
                                          C.byref(o), hip._p(samples), hip._p(acc), C.byref(st)))
        return samples, acc, st.as_dict()


def laplace_inverse(ctx, deck, out_dir=None, resolve=False, chains=1, seed=0, steps=None, opts=None):
    """The laplaceInverse application (del2/apps/reference/laplaceInverse/laplaceInverse.cpp:62-170)
    on its own input deck: input.dat, the parameter map, the inference input,
    the SMF mesh, GP anchor coordinates, phiMat, sensor coordinates and
    readings. Runs hyperParamLearning's chain (one per copy when chains > 1,
    all in lockstep), discards the first 30 % of each chain (burn-in,
    hyperParamLearning.hpp:228-231), and returns the posterior means of the
    natural hyperparameters with the retained samples. With out_dir, writes
    hpsMeanFile and sampleMCMCFile as the reference does (first chain)."""
    def path(name):
        return os.path.join(deck, os.path.basename(name))

    top = io.read_properties(path("input.dat"))
    inf = io.read_properties(path(top["infInputFile"]))
    hp = io.read_param_map(path(top["allHyperFile"]))
    mesh = io.read_smf(path(top["meshFile"]))
    anchors = io.read_matrix_data(path(top["conductCoordFile"]))
    cols, vals = io.read_matrix_phi(path(inf["phiMatFile"]))
    xs = io.read_coord_data(path(inf["sensorPosFile"]), mesh.dim)
    Y = io.read_matrix_data(path(inf["sensorValFile"]))
    # homogeneous Dirichlet at x = 0 and x = 1 (fem.hpp:31-43)
    x0 = mesh.coords[:, 0]
    bc = np.nonzero((np.abs(x0) < 1e-12) | (np.abs(x0 - 1.0) < 1e-12))[0].astype(np.int32)
    na = anchors.shape[0]
    ids = [f"gaussianProcess.w{i}" for i in range(na)] + ["mismatch.sensorEps"]
    gp = (hp["gaussianProcess.sigma"], hp["gaussianProcess.length"], hp["gaussianProcess.eps"])
    # laplaceInverse uses LINE2 with three quadrature points (laplaceInverse.cpp:84-90)
    fg = ForwardGaussLHSHip(ctx, mesh, bc, anchors, gp, cols, vals, copies=chains, nq1d=3, opts=opts)
    pm = np.array([np.log(hp["priorMean." + k]) for k in ids[:-1]] + [hp["priorMean." + ids[-1]]])
    pv = np.array([hp["priorCov." + k] for k in ids])
    prior = int(inf.get("enrichWithPrior", 1)) != 0
    fg.set_data(Y, xs, rho=hp["rho"], mis_sigma=hp["mismatch.sigma"], mis_ell=hp["mismatch.length"],
                prior_mu=pm if prior else None, prior_var=pv if prior else None)
    # the FE solution at the map's own hyperparameters (inference.hpp:138-139)
    w0 = np.array([np.log(hp[k]) for k in ids[:-1]] + [np.log(hp["mismatch.sensorEps"])])
    fg.updateSolution(w0)
    fg.fix_mean(0)
    theta0 = np.array([np.log(hp["init." + k]) for k in ids])
    propsd = np.array([hp["propStDev." + k] for k in ids])
    n = int(steps if steps is not None else inf["numSamples"])
    samples, acc, st = fg.mcmc(theta0, propsd, n, seed=seed, resolve=resolve, opts=opts)
    b = int(n * 3.0 / 10.0)
    kept = samples[:, b:, :]
    nat = np.exp(kept[:, :, :-1])
    means = nat.reshape(-1, na + 1).mean(axis=0)
    if out_dir is not None:
        io.write_param_means(os.path.join(out_dir, os.path.basename(inf["hpsMeanFile"])), ids, means)
        io.write_samples(os.path.join(out_dir, os.path.basename(inf["sampleMCMCFile"])), nat[0], kept[0, :, -1])
    fg.close()
    return {"ids": ids, "means": means, "samples": kept, "acceptance": acc, "solve": st}
