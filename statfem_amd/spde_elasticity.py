"""statFEM for 3D linear elasticity with an SPDE Matern body-force field.

Forward model: A_e u = F_bar + B g with A_e the hex8 elasticity operator
(solid/fem/ElementStatic.ipp:284-321) in row-zero form, F_bar the mean body
force (del2/fem/ElementPotential.ipp:146-175 on each component), B the lumped
load of a nodal scalar field on the vertical component (zero on clamped
DOFs, as CovarianceForcing zeroes covVec on constrained DOFs,
statX/infer/CovarianceForcing.hpp:101-114), and g ~ N(0, Q^-1) the Matern
field of the SPDE approach with alpha = 2 (beta = 1):

    Q = tau^2 kappa^4 A M^-1 A,  A = M_L + G / kappa^2
    (statX/inferQ/PrecisionMatrixRational.ipp:24-39,177-209,258-298)

so Q^-1 = (tau kappa^2)^-2 A^-1 M_L A^-1. With R = P A_e^-1 (one transposed
multi-RHS elasticity solve), the prior at the sensors is

    C_u = R B Q^-1 B^T R^T = (tau kappa^2)^-2 Y^T M_L Y,  Y = A^-1 (B^T R^T)

(one multi-RHS Matern solve and a mass-weighted Gram product), the same
structure as ForwardGaussRHS::updateSolution (statX/infer/ForwardGaussRHS.hpp:95-139)
with the dense C_f replaced by the sparse SPDE precision. Conditioning on the
readings follows LikelihoodKOH (statX/infer/LikelihoodKOH.ipp:310-361,441-506)
with no model discrepancy (sigma_d = 0), so K_s = sigma_e^2 I."""
import time

import numpy as np

from . import hip


class SPDEElasticity:
    def __init__(self, ctx, cfg, opts=None, opts_matern=None):
        self.ctx = ctx
        self.cfg = cfg
        m = cfg["mesh"]
        self.opts = opts or hip.default_opts(tol=1e-10)
        self.opts_m = opts_matern or hip.default_opts(tol=1e-12)
        self.times = {}
        t0 = time.perf_counter()
        # elasticity operator, mean load, clamped face
        # both operators live on the mesh's tensor grid: declared, so their
        # solves run the 3D stencil layout with the multigrid preconditioner
        self.se = hip.SystemSolveHip(ctx, m, dofpn=3)
        self.se.set_grid()
        self.se.assemble(hip.ELASTICITY, cfg["E"], cfg["nu"], nq1d=cfg["nq1d"])
        self.se.assemble_rhs(nq1d=cfg["nq1d"], fconst=[0.0, 0.0, cfg["fbar"]])
        self.se.applyConstraints(cfg["bc"])
        # scalar Matern operator A = M_L + G / kappa^2 on the same mesh
        kap = cfg["kappa"]
        self.sm = hip.SystemSolveHip(ctx, m, dofpn=1)
        self.sm.set_grid()
        self.sm.assemble(hip.MATERN, kap, 1.0, nq1d=cfg["nq1d"], factor=1.0 / kap ** 2)
        self.mass = self.sm.lumped_mass(cfg["nq1d"])
        tau = hip.matern_tau(cfg["sigma_g"], kap, cfg["nu_m"], 3)
        self.gscale = 1.0 / (tau * kap ** 2)
        w = self.mass.download()
        w[cfg["bottom"]] = 0.0
        self.w_load = ctx.to_device(w)
        self.times["assemble"] = time.perf_counter() - t0
        self.ndof = self.se.ndof
        self.ns = len(cfg["cols"])

    def _Pt(self):
        cfg = self.cfg
        Pt = np.zeros((self.ndof, self.ns))
        for s in range(self.ns):
            np.add.at(Pt[:, s], cfg["cols"][s], cfg["vals"][s])
        return Pt

    def prior(self):
        """Mean (ns) and covariance (ns x ns) of the vertical displacement at the
        sensors, plus the device blocks the posterior field needs."""
        ctx, cfg = self.ctx, self.cfg
        t0 = time.perf_counter()
        Pt = ctx.to_device(self._Pt())
        Rt, st1 = self.se.solve(Pt, trans=True, opts=self.opts)            # N_e x ns
        W = hip.dof_gather(ctx, self.w_load, Rt, 3, 2)                       # B^T R^T
        Y, st2 = self.sm.solve(W, opts=self.opts_m)                          # A^-1 B^T R^T
        MY = hip.row_scale(ctx, self.mass, Y)
        Cu = hip.gram(ctx, Y, MY) * self.gscale ** 2
        ubar, st3 = self.se.solve(self._F_device(), opts=self.opts)
        mean = self._project(ubar.download()[:, 0])
        self.times["prior"] = time.perf_counter() - t0
        self._Y, self._MY, self._ubar = Y, MY, ubar
        self.stats = {"Rt": st1, "matern": st2, "mean": st3}
        return mean, Cu

    def _F_device(self):
        F = self.se.rhs()
        return self.ctx.to_device(F.reshape(-1, 1))

    def _project(self, u):
        cfg = self.cfg
        return np.array([np.dot(cfg["vals"][s], u[cfg["cols"][s]]) for s in range(self.ns)])

    def synthesize(self, seed=None):
        """One reading of the vertical displacement at the sensors for a seeded
        Matern load sample g = (tau kappa^2)^-1 A^-1 M_L^{1/2} z (FP64)."""
        ctx, cfg = self.ctx, self.cfg
        rng = np.random.default_rng(cfg["noise_seed"] if seed is None else seed)
        z = rng.standard_normal(self.sm.ndof)
        rhs = np.sqrt(self.mass.download()) * z
        g, _ = self.sm.solve(rhs.reshape(-1, 1), opts=hip.default_opts(tol=1e-12))
        g = g.download()[:, 0] * self.gscale
        Fg = hip.dof_scatter(ctx, self.w_load, ctx.to_device(g.reshape(-1, 1)), 3, 2).download()[:, 0]
        F = self.se.rhs() + Fg
        u, _ = self.se.solve(F.reshape(-1, 1), opts=hip.default_opts(tol=1e-12))
        u = u.download()[:, 0]
        y = self._project(u) + cfg["sigma_e"] * rng.standard_normal(self.ns)
        return y, g, u

    def posterior(self, mean, Cu, y, field=False):
        """Sensor-space posterior mean and variance, log det(C_u + sigma_e^2 I)
        and, with field=True, the posterior mean displacement on every DOF."""
        ctx, cfg = self.ctx, self.cfg
        t0 = time.perf_counter()
        lik = hip.LikelihoodKOHHip(ctx, cfg["xs"], Cu, mean, np.atleast_2d(y), cfg["sigma_e"])
        pm, pc = lik.giveProjFEconditionedData(1.0, 0.0, 1.0)
        ld, st = lik.giveLogDeterminant([[1.0, 0.0, 1.0]])
        out = {"mean": pm, "var": np.diag(pc).copy(), "logdet": float(ld[0]), "status": int(st[0])}
        if field:
            # C_u P^T over the field: A_e^-1 B Q^-1 B^T R^T = (tau kappa^2)^-2 A_e^-1 B A^-1 M_L Y
            Z, _ = self.sm.solve(self._MY, opts=self.opts_m)
            Fz = hip.dof_scatter(ctx, self.w_load, Z, 3, 2)
            Fld, st4 = self.se.solve(Fz, opts=self.opts)
            Fld = hip.row_scale(ctx, ctx.to_device(np.full(self.ndof, self.gscale ** 2)), Fld, out=Fld)
            self._field = Fld
            out["field_mean"] = lik.giveFEconditionedDataWoodbury(1.0, 0.0, 1.0, Fld, self._ubar).download()
            self.stats["field"] = st4
        self.times["posterior"] = time.perf_counter() - t0
        return out
