"""Many independent statFEM problems of uneven size in one system.

The meshes are joined into a disjoint union (block-diagonal operator), assembled
and constrained once, and the multi-RHS transposed solve of every problem runs
in the same kernel launches (`sfh_solve_batched`: per-problem, per-column PCG
scalars and convergence). Each problem's forcing covariance is applied on its
own tensor grid and its C_u = R C_f R^T comes from a per-problem Gram product,
so the result per problem is ForwardGaussRHS::updateSolution
(statX/infer/ForwardGaussRHS.hpp:95-139) on that problem alone."""
import ctypes as C

import numpy as np

from . import hip
from . import mesh as M


class BatchedForwardGauss:
    def __init__(self, ctx, problems, nq1d=2):
        self.ctx = ctx
        self.probs = problems
        meshes = [p["mesh"] for p in problems]
        self.mesh, off = M.union(meshes)
        self.seg = np.ascontiguousarray(off, dtype=np.int32)
        self.nseg = len(problems)
        # the widest problem's sensors, rounded up to an even count: the
        # right-hand-side blocks then have an even row stride, which the dense
        # products load as 16-byte pairs (the extra column is zero and the
        # solver skips it)
        self.kmax = (max(len(p["cols"]) for p in problems) + 1) // 2 * 2
        s = hip.SystemSolveHip(ctx, self.mesh)
        s.assemble(hip.POISSON, 1.0, nq1d=nq1d)
        s.assemble_rhs(nq1d=nq1d, fconst=1.0)
        bc = np.concatenate([p["bc"] + off[q] for q, p in enumerate(problems)]).astype(np.int32)
        s.applyConstraints(bc)
        nx = np.array([len(p["mesh"].grid[0]) for p in problems], np.int32)
        ny = np.array([len(p["mesh"].grid[1]) for p in problems], np.int32)
        hip._check(hip.lib().sfh_system_set_segment_grids(s.h, self.nseg, hip._p(self.seg), hip._p(nx), hip._p(ny)))
        self.sys = s
        cv = s.lumped_mass(nq1d).download()
        cv[bc] = 0.0
        self.cv = ctx.to_device(cv)
        N = self.mesh.nnodes
        Pt = np.zeros((N, self.kmax))
        for q, p in enumerate(problems):
            for j in range(len(p["cols"])):
                np.add.at(Pt[:, j], p["cols"][j] + off[q], p["vals"][j])
        self.Pt = ctx.to_device(Pt)
        self.F = ctx.to_device(s.rhs().reshape(-1, 1))
        # per-problem forcing covariance: grid shapes, axes (x then y, problem
        # after problem), sigma_f and ell_f, for one batched C_f call
        self.nx = np.array([len(p["mesh"].grid[0]) for p in problems], np.int32)
        self.ny = np.array([len(p["mesh"].grid[1]) for p in problems], np.int32)
        self.axes = np.concatenate([np.r_[p["mesh"].grid[0], p["mesh"].grid[1]] for p in problems]).astype(np.float64)
        self.sig = np.array([p["sigma_f"] for p in problems], np.float64)
        self.ell = np.array([p["ell_f"] for p in problems], np.float64)
        self.PPt = [_sparse_ppt(p["cols"], p["vals"]) for p in problems]

    def _buffers(self):
        """R^T, V = C_f R^T, the Gram blocks and the means, allocated on the
        first update and kept for the following ones."""
        if getattr(self, "_buf", None) is None:
            ctx, k, N = self.ctx, self.kmax, self.mesh.nnodes
            self._buf = (ctx.empty((N, k)), ctx.empty((N, k)), ctx.empty((self.nseg, k, k)),
                         ctx.empty((self.nseg, k, 1)))
        return self._buf

    def updateSolution(self, opts=None):
        """Prior mean and covariance at each problem's sensors."""
        ctx, k = self.ctx, self.kmax
        N = self.mesh.nnodes
        Rt, V, G, Mn = self._buffers()
        st = hip.SolveStats()
        o = opts or hip.default_opts()
        hip._check(hip.lib().sfh_solve_batched(self.sys.h, self.nseg, hip._p(self.seg), 1, k, self.Pt.ptr, Rt.ptr,
                                               C.byref(o), C.byref(st)))
        hip._check(hip.lib().sfh_covforcing_batched(ctx.h, self.nseg, hip._p(self.seg), hip._p(self.nx),
                                                    hip._p(self.ny), hip._p(self.axes), self.cv.ptr,
                                                    hip._p(self.sig), hip._p(self.ell), k, Rt.ptr, V.ptr))
        hip._check(hip.lib().sfh_gram_segmented(ctx.h, self.nseg, hip._p(self.seg), k, k, Rt.ptr, k, V.ptr, k, G.ptr))
        hip._check(hip.lib().sfh_gram_segmented(ctx.h, self.nseg, hip._p(self.seg), k, 1, Rt.ptr, k, self.F.ptr, 1,
                                                Mn.ptr))
        G, Mn = G.download(), Mn.download()   # (host copies; the device blocks stay for the next update)
        means, Cus = [], []
        for q, p in enumerate(self.probs):
            ns = len(p["cols"])
            reg = (p["sigma_f"] * p["reg"]) ** 2
            Cus.append(G[q, :ns, :ns] + reg * self.PPt[q])
            means.append(Mn[q, :ns, 0].copy())
        self.Rt = Rt   # R^T of the last update (overwritten by the next one)
        return means, Cus, st.as_dict()


def _sparse_ppt(cols, vals):
    """P P^T of sensor rows given as (cols, vals) of shape (ns, nen)."""
    ns = cols.shape[0]
    out = np.zeros((ns, ns))
    rows = [dict() for _ in range(ns)]
    for i in range(ns):
        for c, v in zip(cols[i], vals[i]):
            rows[i][int(c)] = rows[i].get(int(c), 0.0) + float(v)
    for i in range(ns):
        for j in range(i, ns):
            d = sum(v * rows[j].get(c, 0.0) for c, v in rows[i].items())
            out[i, j] = out[j, i] = d
    return out


class SequentialForwardGauss:
    """The same priors one problem at a time: one system, one ForwardGaussRHS
    update and its own solves per problem."""

    def __init__(self, ctx, problems, nq1d=2):
        self.probs = problems
        self.fgs = []
        for p in problems:
            s = hip.SystemSolveHip(ctx, p["mesh"])
            s.assemble(hip.POISSON, 1.0, nq1d=nq1d)
            s.assemble_rhs(nq1d=nq1d, fconst=1.0)
            s.applyConstraints(p["bc"])
            s.set_grid()
            self.fgs.append(hip.ForwardGaussRHSHip(s, p["cols"], p["vals"], reg_eps=p["reg"]))

    def updateSolution(self, opts=None):
        means, Cus, iters = [], [], 0
        for p, fg in zip(self.probs, self.fgs):
            m, Cu, st = fg.updateSolution(p["sigma_f"], p["ell_f"], opts=opts)
            means.append(m)
            Cus.append(Cu)
            iters += st["inner_iters"]
        return means, Cus, iters

    def close(self):
        for fg in self.fgs:
            fg.sys.close()
