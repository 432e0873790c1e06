"""FP64 oracle of ForwardGaussRHS::updateSolution on large 2D tensor grids
(statX/infer/ForwardGaussRHS.hpp:95-139 with CovarianceForcing,
statX/infer/CovarianceForcing.ipp:37-109), composed from the oracle's C++
pieces. Test infrastructure: used by the test suite and the CPU-baseline leg
of the benchmark only.

At these sizes the reference's dense O(N^2) C_f is out of reach, so C_f W is
formed through the Kronecker structure of the squared-exponential kernel on a
tensor grid, K = sigma^2 K_y (x) K_x, and every call checks that product on
sampled rows against direct N-term sums of the reference's C_f entries
(or_covforcing_rows). Solves use the banded Cholesky factor of the
symmetrically eliminated operator, which gives the row-zero operator's
free-DOF solution for homogeneous constraints; the constrained components of
transposed solutions are restored as r_c = p_c - sum_f A[f][c] r_f."""
import numpy as np

from . import oracle as orc


def separable_cf(grid, covvec, sigma, ell, W):
    """C_f W on a 2D tensor grid (nodes x fastest)"""
    gx, gy = grid
    kx = np.exp(-(gx[:, None] - gx[None, :]) ** 2 / (2 * ell * ell))
    ky = np.exp(-(gy[:, None] - gy[None, :]) ** 2 / (2 * ell * ell))
    out = np.empty_like(W)
    for c in range(W.shape[1]):
        G = (covvec * W[:, c]).reshape(len(gy), len(gx))
        out[:, c] = covvec * (sigma * sigma * (ky @ G @ kx.T)).ravel()
    return out


class ForwardGauss2D:
    """The factorised operator and forcing weights of one scalar 2D problem."""

    def __init__(self, mesh, bc, kappa, nq1d=2, threads=16):
        import scipy.sparse as sp
        self.mesh, self.bc, self.threads = mesh, bc, threads
        N = mesh.nnodes
        F = orc.body_force(mesh, nq1d=nq1d, fconst=1.0)
        self.rp, self.ci, self.va, self.F = orc.assemble(mesh, orc.POISSON, kappa, nq1d=nq1d, bc_dofs=bc, F=F)
        self.A = sp.csr_matrix((self.va, self.ci, self.rp), shape=(N, N))
        mask = np.zeros(N, np.uint8)
        mask[bc] = 1
        self.ch = orc.Chol(self.rp, self.ci, self.va, mask)
        self.cv = orc.body_force(mesh, nq1d=nq1d, fconst=1.0)
        self.cv[bc] = 0.0

    def solve(self, B):
        return self.ch.solve(B, threads=self.threads)

    def transposed(self, P):
        """A^-T P for the row-zero operator"""
        Pf = P.copy()
        Pf[self.bc] = 0.0
        R = self.solve(Pf)
        R[self.bc] = 0.0
        R[self.bc] = P[self.bc] - (self.A.T @ R)[self.bc]
        return R

    def cf(self, sigma, ell, W, check_rows=32, seed=0):
        """C_f W, spot-checked against the direct sums on `check_rows` rows"""
        V = separable_cf(self.mesh.grid, self.cv, sigma, ell, W)
        nz = np.flatnonzero(self.cv)
        if check_rows and len(nz):
            rows = np.sort(np.random.default_rng(seed).choice(nz, size=min(check_rows, len(nz)), replace=False))
            Vd = orc.covforcing_rows(self.mesh.coords, self.cv, sigma, ell, W, rows.astype(np.int32))
            scale = np.abs(V).max()
            if np.abs(Vd - V[rows]).max() > 1e-12 * scale:
                raise AssertionError("separable C_f product differs from the direct sums")
        return V

    def forward_gauss(self, cols, vals, sigma, ell, reg, sample=None):
        """Projected mean (all sensors) and C_u on the sampled sensors (all if
        sample is None), with R^T on those columns."""
        N = self.mesh.nnodes
        P = np.zeros((N, len(cols)))
        for s in range(len(cols)):
            np.add.at(P[:, s], np.asarray(cols[s]), np.asarray(vals[s]))
        ubar = self.solve(self.F)
        mean = P.T @ ubar
        idx = np.arange(P.shape[1]) if sample is None else np.asarray(sample)
        Ps = P[:, idx]
        R = self.transposed(Ps)
        V = self.cf(sigma, ell, R)
        Cu = R.T @ V + (sigma * reg) ** 2 * (Ps.T @ Ps)
        return dict(mean=mean, Cu=Cu, Rt=R, V=V, ubar=ubar, sample=idx, P=P)


def elasticity_spde_reference(cfg, y):
    """FP64 oracle of the elasticity statFEM at a size the banded Cholesky
    factorisations cover: prior mean and C_u at the sensors, the field block
    C_u P^T, and the sensor posterior (mean, variance, log det) for reading y.
    Both operators are SPD after symmetric elimination of the clamped face
    (the row-zero operator's free-DOF solution for homogeneous constraints).
    The prior is C_u = R B Q^-1 B^T R^T with Q^-1 = (tau kappa^2)^-2 A^-1 M_L A^-1
    (statX/inferQ/PrecisionMatrixRational.ipp:24-39,177-209,258-298), the
    posterior that of LikelihoodKOH with K_s = sigma_e^2 I
    (statX/infer/LikelihoodKOH.ipp:310-361,441-506)."""
    m = cfg["mesh"]
    N = 3 * m.nnodes
    F = orc.body_force(m, nq1d=2, dofpn=3, fconst=[0.0, 0.0, cfg["fbar"]])
    rp, ci, va, F = orc.assemble(m, orc.ELASTICITY, cfg["E"], cfg["nu"], nq1d=2, bc_dofs=cfg["bc"], F=F)
    mask = np.zeros(N, np.uint8)
    mask[cfg["bc"]] = 1
    ch = orc.Chol(rp, ci, va, mask)
    kap = cfg["kappa"]
    rpm, cim, vam, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    chm = orc.Chol(rpm, cim, vam, np.zeros(m.nnodes, np.uint8))
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    w = mass.copy()
    w[cfg["bottom"]] = 0.0
    gscale = 1.0 / (orc.matern_tau(cfg["sigma_g"], kap, cfg["nu_m"], 3) * kap ** 2)
    ns = len(cfg["cols"])
    P = np.zeros((N, ns))
    for s in range(ns):
        np.add.at(P[:, s], cfg["cols"][s], cfg["vals"][s])
    R = ch.solve(P, threads=16)                       # free rows of A_e^-T P^T (B^T ignores the clamped face)
    Y = chm.solve(w[:, None] * R[2::3], threads=16)   # A^-1 B^T R^T
    Cu = gscale ** 2 * Y.T @ (mass[:, None] * Y)
    ubar = ch.solve(F)
    mean = P.T @ ubar
    Z = chm.solve(mass[:, None] * Y, threads=16)
    BZ = np.zeros((N, ns))
    BZ[2::3] = w[:, None] * Z
    field = gscale ** 2 * ch.solve(BZ, threads=16)    # C_u P^T = A_e^-1 B Q^-1 B^T R^T
    Ks = np.eye(ns) * cfg["sigma_e"] ** 2
    pm, pc = orc.post_proj(Cu, mean, np.atleast_2d(y), Ks, 1.0)
    sign, logdet = np.linalg.slogdet(Cu + Ks)
    S = Cu + Ks
    fmean = ubar + field @ np.linalg.solve(S, y - mean)
    return dict(ch=ch, chm=chm, w=w, gscale=gscale, F=F, Cu=Cu, mean=mean, field=field, ubar=ubar,
                post_mean=pm, post_var=np.diag(pc).copy(), logdet=logdet, sign=sign, field_mean=fmean)
