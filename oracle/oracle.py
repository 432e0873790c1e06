"""ctypes bindings of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, the smoke check and the CPU
baseline leg of bench.py. The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

POISSON, ELASTICITY, MATERN = 0, 1, 2

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def _dpn():
    return C.c_void_p


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    i, d, v, ll = C.c_int, C.c_double, C.c_void_p, C.c_longlong
    sig = {
        "or_quad_rule": (i, [i, i, _dp, _dp]),
        "or_element_matrices": (i, [i, i, i, i, _dp, i, _ip, _dp, _dp, d, _dp]),
        "or_body_force": (i, [i, i, i, _dp, i, _ip, i, v, v, d, _dp]),
        "or_triplet_new": (v, [i]),
        "or_triplet_free": (None, [v]),
        "or_triplet_insert_elements": (i, [v, i, i, _ip, _dp]),
        "or_triplet_apply_constraints": (i, [v, v, i, _ip, v, d, d]),
        "or_triplet_nnz": (ll, [v]),
        "or_triplet_csr": (i, [v, _ip, _ip, _dp]),
        "or_lu_factor": (v, [i, _ip, _ip, _dp, C.POINTER(C.c_int)]),
        "or_lu_free": (None, [v]),
        "or_lu_solve": (i, [v, i, i, _dp, _dp]),
        "or_chol_factor": (v, [i, _ip, _ip, _dp, v, C.POINTER(C.c_int)]),
        "or_chol_free": (None, [v]),
        "or_chol_solve": (i, [v, i, _dp, _dp]),
        "or_chol_solve_mt": (i, [v, i, _dp, _dp, i]),
        "or_spmv": (i, [i, _ip, _ip, _dp, i, _dp, _dp]),
        "or_philox4x32_10": (None, [_up, _up, _up]),
        "or_mcmc_normals": (i, [C.c_uint64, i, i, _dp, _dp]),
        "or_lhs_draws": (i, [C.c_uint64, i, i, i, _dp, _dp]),
        "or_se_kernel": (d, [i, _dp, _dp, d, d]),
        "or_covforcing_project": (i, [i, i, _dp, _dp, d, d, d, i, _dp, _dp]),
        "or_covforcing_apply": (i, [i, i, _dp, _dp, d, d, i, _dp, _dp]),
        "or_covforcing_rows": (i, [i, i, _dp, _dp, d, d, i, _dp, i, _ip, _dp]),
        "or_mismatch_cov": (i, [i, i, _dp, d, d, d, _dp]),
        "or_loglik": (i, [i, i, _dp, _dp, _dp, _dp, d, C.POINTER(C.c_double)]),
        "or_post_proj": (i, [i, i, _dp, _dp, _dp, _dp, d, _dp, _dp]),
        "or_post_woodbury": (i, [i, i, i, _dp, _dp, _dp, _dp, _dp, d, _dp, v]),
        "or_gp": (i, [i, i, _dp, _dp, d, d, d, i, _dp, C.POINTER(C.c_double), _dp, _dp]),
        "or_gp_cov": (i, [i, i, _dp, d, d, d, i, _dp, _dp]),
        "or_mcmc_koh": (i, [i, i, i, _dp, _dp, _dp, _dp, d, _dp, _dp, i, C.c_uint64, i, i,
                            _dp, _dp, _dp, C.POINTER(C.c_double)]),
        "or_mcmc_gauss": (i, [d, d, d, d, i, C.c_uint64, i, _dp, C.POINTER(C.c_double)]),
        "or_typical_readings": (i, [i, i, i, _dp, _dp, _dp, d, d, d, d, C.c_uint64, _dp]),
        "or_matern_tau": (d, [d, d, d, i]),
        "or_precision_apply": (i, [i, _ip, _ip, _dp, _dp, d, d, d, i, i, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ----------------------------------------------------------------- FE half
def quad_rule(kind, nq1d):
    w = np.zeros(27)
    xi = np.zeros(81)
    n = lib().or_quad_rule(kind, nq1d, w, xi)
    return w[:n], xi[: 3 * n].reshape(n, 3)


def element_matrices(mesh, physics, p1, p2=None, nq1d=2, factor=1.0):
    dofpn = 3 if physics == ELASTICITY else 1
    ned = mesh.nen * dofpn
    Ke = np.zeros((mesh.nel, ned, ned))
    p1 = _c(np.broadcast_to(p1, (mesh.nel,)))
    p2 = _c(np.broadcast_to(0.0 if p2 is None else p2, (mesh.nel,)))
    rc = lib().or_element_matrices(physics, mesh.kind, nq1d, mesh.nnodes, _c(mesh.coords),
                                   mesh.nel, _c(mesh.conn, np.int32), p1, p2, factor, Ke)
    if rc != 0:
        raise RuntimeError(f"or_element_matrices: {rc}")
    return Ke


def body_force(mesh, nq1d=2, dofpn=1, fconst=None, fnodal=None, factor=1.0):
    F = np.zeros(mesh.nnodes * dofpn)
    fc = None if fconst is None else _c(np.broadcast_to(fconst, (dofpn,)))
    fn = None if fnodal is None else _c(fnodal)
    rc = lib().or_body_force(mesh.kind, nq1d, mesh.nnodes, _c(mesh.coords), mesh.nel,
                             _c(mesh.conn, np.int32), dofpn, _ptr(fc), _ptr(fn), factor, F)
    if rc != 0:
        raise RuntimeError(f"or_body_force: {rc}")
    return F


class TripletSystem:
    """Triplet-backed assembly with the reference's sort-merge and row-zero
    Dirichlet semantics (corlib/Triplet.hpp, corlib/SystemSolveEigenSparse.hpp)."""

    def __init__(self, ndof):
        self.n = ndof
        self.h = lib().or_triplet_new(ndof)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_triplet_free(self.h)
            self.h = None

    def insert_elements(self, edofs, Ke):
        edofs = _c(edofs, np.int32)
        Ke = _c(Ke)
        lib().or_triplet_insert_elements(self.h, edofs.shape[0], edofs.shape[1], edofs, Ke)

    def apply_constraints(self, F, dofs, values=None, factor=1.0, scalar=1.0):
        dofs = _c(dofs, np.int32)
        vals = None if values is None else _c(values)
        lib().or_triplet_apply_constraints(self.h, _ptr(F), len(dofs), dofs, _ptr(vals), factor, scalar)

    def csr(self):
        nnz = lib().or_triplet_nnz(self.h)
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        va = np.zeros(nnz)
        rc = lib().or_triplet_csr(self.h, rp, ci, va)
        if rc != 0:
            raise RuntimeError("or_triplet_csr: row count mismatch")
        return rp, ci, va


def assemble(mesh, physics, p1, p2=None, nq1d=2, factor=1.0, bc_dofs=None, bc_vals=None, F=None):
    """Full reference-form assembly: element matrices -> Triplet -> constraints -> CSR."""
    dofpn = 3 if physics == ELASTICITY else 1
    Ke = element_matrices(mesh, physics, p1, p2, nq1d, factor)
    ts = TripletSystem(mesh.nnodes * dofpn)
    ts.insert_elements(mesh.element_dofs(dofpn), Ke)
    if F is None:
        F = np.zeros(mesh.nnodes * dofpn)
    else:
        F = np.array(F, dtype=np.float64)
    if bc_dofs is not None and len(bc_dofs):
        ts.apply_constraints(F, bc_dofs, bc_vals)
    rp, ci, va = ts.csr()
    return rp, ci, va, F


class LU:
    """Direct solve with the reference semantics (SparseLU stand-in)."""

    def __init__(self, rp, ci, va):
        st = C.c_int(0)
        self.n = len(rp) - 1
        self.h = lib().or_lu_factor(self.n, _c(rp, np.int32), _c(ci, np.int32), _c(va), C.byref(st))
        if not self.h:
            raise RuntimeError(f"or_lu_factor: {st.value}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_lu_free(self.h)
            self.h = None

    def solve(self, B, trans=False):
        B = _c(B)
        squeeze = B.ndim == 1
        B2 = B.reshape(self.n, -1)
        X = np.zeros_like(B2)
        lib().or_lu_solve(self.h, int(trans), B2.shape[1], _c(B2), X)
        return X.ravel() if squeeze else X


class Chol:
    """Banded Cholesky of the symmetric-elimination operator (CPU baseline)."""

    def __init__(self, rp, ci, va, constrained_mask):
        st = C.c_int(0)
        self.n = len(rp) - 1
        m = _c(constrained_mask, np.uint8)
        self.h = lib().or_chol_factor(self.n, _c(rp, np.int32), _c(ci, np.int32), _c(va),
                                      m.ctypes.data_as(C.c_void_p), C.byref(st))
        if not self.h:
            raise RuntimeError(f"or_chol_factor: {st.value}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_chol_free(self.h)
            self.h = None

    def solve(self, B, threads=1):
        """L L^T X = B; threads > 1 splits the columns over std::threads."""
        B = _c(B)
        squeeze = B.ndim == 1
        B2 = B.reshape(self.n, -1)
        X = np.zeros_like(B2)
        if threads > 1:
            lib().or_chol_solve_mt(self.h, B2.shape[1], _c(B2), X, int(threads))
        else:
            lib().or_chol_solve(self.h, B2.shape[1], _c(B2), X)
        return X.ravel() if squeeze else X


def spmv(rp, ci, va, X):
    X = _c(X)
    squeeze = X.ndim == 1
    n = len(rp) - 1
    X2 = X.reshape(n, -1)
    Y = np.zeros_like(X2)
    lib().or_spmv(n, _c(rp, np.int32), _c(ci, np.int32), _c(va), X2.shape[1], _c(X2), Y)
    return Y.ravel() if squeeze else Y


# --------------------------------------------------------------- statistics
def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_c(ctr, np.uint32), _c(key, np.uint32), out)
    return out


def mcmc_draws(seed, chain, step):
    z = np.zeros(3)
    u = np.zeros(1)
    lib().or_mcmc_normals(seed, chain, step, z, u)
    return z, u[0]


def lhs_draws(seed, chain, step, np_):
    z = np.zeros(np_)
    u = np.zeros(1)
    lib().or_lhs_draws(seed, chain, step, np_, z, u)
    return z, u[0]


def covforcing_project(coords, covvec, sigma, ell, R, cut_eps=0.0):
    coords = _c(coords)
    n, dim = coords.shape
    R = _c(R)
    ns = R.shape[0]
    Cu = np.zeros((ns, ns))
    lib().or_covforcing_project(n, dim, coords, _c(covvec), sigma, ell, cut_eps, ns, R, Cu)
    return Cu


def covforcing_apply(coords, covvec, sigma, ell, W):
    coords = _c(coords)
    n, dim = coords.shape
    W = _c(W).reshape(n, -1)
    V = np.zeros_like(W)
    lib().or_covforcing_apply(n, dim, coords, _c(covvec), sigma, ell, W.shape[1], W, V)
    return V


def covforcing_rows(coords, covvec, sigma, ell, W, rows):
    """rows `rows` of V = C_f W by direct n-term sums (spot checks at large n)"""
    coords = _c(coords)
    n, dim = coords.shape
    W = _c(W).reshape(n, -1)
    rows = _c(rows, np.int32)
    V = np.zeros((len(rows), W.shape[1]))
    rc = lib().or_covforcing_rows(n, dim, coords, _c(covvec), sigma, ell, W.shape[1], W, len(rows), rows, V)
    if rc != 0:
        raise ValueError(f"or_covforcing_rows: {rc}")
    return V


def mismatch_cov(xs, sigma, ell, eps):
    xs = _c(np.atleast_2d(xs))
    if xs.shape[0] == 1 and xs.shape[1] > 1:
        xs = xs.T.copy()
    ns, dim = xs.shape
    Ks = np.zeros((ns, ns))
    lib().or_mismatch_cov(ns, dim, xs, sigma, ell, eps, Ks)
    return Ks


def loglik(Cu, mean, Y, Ks, rho):
    Y = _c(Y)
    no, ns = Y.shape
    out = C.c_double(0.0)
    rc = lib().or_loglik(ns, no, _c(Cu), _c(mean), Y, _c(Ks), rho, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"or_loglik: {rc}")
    return out.value


def post_proj(Cu, mean, Y, Ks, rho):
    Y = _c(Y)
    no, ns = Y.shape
    m = np.zeros(ns)
    c = np.zeros((ns, ns))
    rc = lib().or_post_proj(ns, no, _c(Cu), _c(mean), Y, _c(Ks), rho, m, c)
    if rc != 0:
        raise RuntimeError(f"or_post_proj: {rc}")
    return m, c


def post_woodbury(CuFull, ubar, P, Ks, Y, rho, want_cov=True):
    Y = _c(Y)
    no, ns = Y.shape
    n = CuFull.shape[0]
    m = np.zeros(n)
    c = np.zeros((n, n)) if want_cov else None
    rc = lib().or_post_woodbury(n, ns, no, _c(CuFull), _c(ubar), _c(P), _c(Ks), Y, rho, m, _ptr(c))
    if rc != 0:
        raise RuntimeError(f"or_post_woodbury: {rc}")
    return m, c


def gp(X, y, sigma, ell, eps, Xp):
    X = _c(X).reshape(len(y), -1)
    n, dim = X.shape
    Xp = _c(Xp).reshape(-1, dim)
    ll = C.c_double(0.0)
    mean = np.zeros(Xp.shape[0])
    var = np.zeros(Xp.shape[0])
    rc = lib().or_gp(n, dim, X, _c(y), sigma, ell, eps, Xp.shape[0], Xp, C.byref(ll), mean, var)
    if rc != 0:
        raise RuntimeError(f"or_gp: {rc}")
    return ll.value, mean, var


def gp_cov(X, sigma, ell, eps, Xp, n):
    """predictCovariance at the test locations Xp (np x dim) of the GP on X (n x dim)"""
    X = _c(X).reshape(n, -1)
    dim = X.shape[1]
    Xp = _c(Xp).reshape(-1, dim)
    cov = np.zeros((Xp.shape[0], Xp.shape[0]))
    rc = lib().or_gp_cov(n, dim, X, sigma, ell, eps, Xp.shape[0], Xp, cov)
    if rc != 0:
        raise RuntimeError(f"or_gp_cov: {rc}")
    return cov


def mcmc_koh(Cu, mean, Y, xs, eps, theta0, propsd, nsteps, seed=0, chain=0, prior_mode=0,
             prior_mu=(0, 0, 0), prior_var=(1, 1, 1)):
    Y = _c(Y)
    no, ns = Y.shape
    xs = _c(xs).reshape(ns, -1)
    samples = np.zeros((nsteps, 4))
    acc = C.c_double(0.0)
    lib().or_mcmc_koh(ns, no, xs.shape[1], _c(Cu), _c(mean), Y, xs, eps, _c(theta0), _c(propsd),
                      nsteps, seed, chain, prior_mode, _c(prior_mu), _c(prior_var), samples,
                      C.byref(acc))
    return samples, acc.value


def mcmc_gauss(mu, sd, x0, propsd, nsteps, seed=0, chain=0):
    samples = np.zeros((nsteps, 2))
    acc = C.c_double(0.0)
    lib().or_mcmc_gauss(mu, sd, x0, propsd, nsteps, seed, chain, samples, C.byref(acc))
    return samples, acc.value


def typical_readings(Cu, mean, xs, eps, rho, sd, ld, nr, seed=0):
    ns = len(mean)
    xs = _c(xs).reshape(ns, -1)
    Y = np.zeros((nr, ns))
    rc = lib().or_typical_readings(ns, nr, xs.shape[1], _c(Cu), _c(mean), xs, eps, rho, sd, ld, seed, Y)
    if rc != 0:
        raise RuntimeError(f"or_typical_readings: {rc}")
    return Y


def matern_tau(sigma, kappa, nu, dim):
    return lib().or_matern_tau(sigma, kappa, nu, dim)


def precision_apply(rp, ci, va, mass, sigma, kappa, nu, dim, X):
    X = _c(X)
    n = len(rp) - 1
    X2 = X.reshape(n, -1)
    Y = np.zeros_like(X2)
    lib().or_precision_apply(n, _c(rp, np.int32), _c(ci, np.int32), _c(va), _c(mass), sigma, kappa, nu, dim,
                             X2.shape[1], _c(X2), Y)
    return Y.reshape(X.shape)


# ------------------------------------------- conductivity (LHS) uncertainty
def lhs_conductivity(mesh, anchors, w, gp_sigma, gp_ell, gp_eps):
    """Element conductivities of laplaceInverse (del2/apps/reference/laplaceInverse/
    inference.hpp:72-88): the GaussianProcess mean (statX/infer/GaussianProcess.hpp:117-135)
    at each element centroid given anchor values w."""
    cen = mesh.coords[mesh.conn].mean(axis=1)
    A = np.asarray(anchors, np.float64).reshape(len(w), -1)
    _, kap, _ = gp(A, np.asarray(w, np.float64), gp_sigma, gp_ell, gp_eps, cen.reshape(len(cen), A.shape[1]))
    return kap


def lhs_solution(mesh, bc, anchors, w, gp_sigma, gp_ell, gp_eps, nq1d=2, fconst=1.0):
    """The solve function of laplaceInverse (inference.hpp:90-120, fem.hpp:47-98):
    conductivity, Poisson assembly with the unit body force, row-zero Dirichlet,
    direct solve. Returns (u, kappa)."""
    kap = lhs_conductivity(mesh, anchors, w, gp_sigma, gp_ell, gp_eps)
    F = body_force(mesh, nq1d=nq1d, fconst=fconst)
    rp, ci, va, F = assemble(mesh, POISSON, kap, nq1d=nq1d, bc_dofs=bc, F=F)
    return LU(rp, ci, va).solve(F), kap


def lhs_log_target(theta, mean, Y, xs, rho, mis_sigma, mis_ell, prior_mu=None, prior_var=None):
    """posteriorGivenData (del2/apps/reference/laplaceInverse/hyperParamLearning.hpp:172-190)
    of one set theta = (log w_0 .. log w_{na-1}, log sensorEps) given the projected
    mean: the KOH log-likelihood with C_u = 0 (ForwardGaussLHS.hpp:71-79), plus
    theta[na], plus (prior_mu given) the log normal density of
    (theta[:na], exp(theta[na])) with diagonal covariance prior_var."""
    theta = np.asarray(theta, np.float64)
    na = len(theta) - 1
    eps = np.exp(theta[na])
    ns = len(mean)
    Ks = mismatch_cov(np.asarray(xs, np.float64).reshape(ns, -1), mis_sigma, mis_ell, eps)
    ll = loglik(np.zeros((ns, ns)), np.asarray(mean, np.float64), Y, Ks, rho)
    y = ll + theta[na]
    if prior_mu is not None:
        hp = np.r_[theta[:na], eps]
        pv = np.asarray(prior_var, np.float64)
        y += -0.5 * np.sum((hp - np.asarray(prior_mu)) ** 2 / pv) - 0.5 * (len(hp) * np.log(2 * np.pi) + np.log(pv).sum())
    return y


def lhs_chain(theta0, propsd, steps, seed, chain, target):
    """MCMC<np>::sample (statX/infer/MCMC.hpp:123-158) with the Philox draws of
    or_lhs_draws; target(theta) -> log target. Returns (steps x (np + 1)
    states and targets, acceptance ratio)."""
    x = np.array(theta0, np.float64)
    y = target(x)
    out = np.zeros((steps, len(x) + 1))
    acc = 0
    for t in range(steps):
        z, u = lhs_draws(seed, chain, t, len(x))
        xp = x + np.asarray(propsd) * z
        yp = target(xp)
        if np.log(u) <= min(0.0, yp - y):
            x, y = xp, yp
            acc += 1
        out[t, :-1] = x
        out[t, -1] = y
    return out, acc / max(steps, 1)


# ------------------------------------- sparse-precision GP (GaussianProcessQ)
def gpq_precision_dense(rp, ci, va, mass, sigma, kappa, nu, dim):
    """The SPDE precision Q of PrecisionMatrixRational::givePrecisionMatrix for
    integer beta (statX/inferQ/PrecisionMatrixRational.ipp:24-39,177-209,278-298)
    as a dense matrix: precision_apply on the identity."""
    n = len(rp) - 1
    return precision_apply(rp, ci, va, mass, sigma, kappa, nu, dim, np.eye(n))


def gpq_frac_factor(rp, ci, va, mass, roots):
    """roots[0] prod_j (I - roots[j] M^-1 A), dense: P_l of giveLeftFracMatrix
    (statX/inferQ/PrecisionMatrixRational.ipp:41-98) for roots = bot.out's
    (b, d_0 ..), P_r of giveRightFracMatrix (:100-161) for top.out's (a, c_0 ..)."""
    n = len(rp) - 1
    T = spmv(rp, ci, va, np.eye(n)) / np.asarray(mass)[:, None]
    F = np.eye(n)
    for r in roots[1:]:
        F = F @ (np.eye(n) - r * T)
    return roots[0] * F


def gpq_precision_frac_dense(rp, ci, va, mass, sigma, kappa, nu, dim, top, bot):
    """givePrecisionMatrix (PrecisionMatrixRational.ipp:24-39) for a fractional
    beta, Q_v = P_l^T Q P_l with Q from the integer part alpha = max(1, floor beta)
    (:177-209), and giveRightFracMatrix's P_r. Returns (Q_v, P_r)."""
    Q = gpq_precision_dense(rp, ci, va, mass, sigma, kappa, nu, dim)
    Pl = gpq_frac_factor(rp, ci, va, mass, bot)
    return Pl.T @ Q @ Pl, gpq_frac_factor(rp, ci, va, mass, top)


def gpq_loglik(Q, P, Y, eps, Pr=None):
    """GaussianProcessQ::giveLogLikelihood (statX/inferQ/GaussianProcessQ.ipp:37-99),
    dense Cholesky factors in place of the sparse LLT, with the effective
    projection P_v = P P_r (:54; P_r = I for integer beta). Q: N x N, P: ns x N,
    Y: ns x no (the reference's obsValues layout)."""
    ns, no = Y.shape
    N = Q.shape[0]
    if Pr is not None:
        P = P @ Pr
    Am = eps * eps * Q + no * P.T @ P                                  # step 1a
    L = np.linalg.cholesky(Am)
    s = Y.sum(axis=1)
    fixed = P @ np.linalg.solve(L.T, np.linalg.solve(L, P.T @ s))
    tmp1 = -0.5 * sum(Y[:, i] @ ((Y[:, i] - fixed) / (eps * eps)) for i in range(no))   # step 1b
    tmp2a = 2.0 * np.log(np.diag(L)).sum()                            # step 2a
    tmp2b = 2.0 * np.log(np.diag(np.linalg.cholesky(Q))).sum()        # step 2b
    diff = ns * no - N                                                 # step 2c
    tmp2 = -0.5 * (diff * np.log(eps * eps) + tmp2a - tmp2b)
    tmp3 = -0.5 * no * ns * np.log(2.0 * np.pi)                       # step 3
    return tmp1 + tmp2 + tmp3


def gpq_predict(Q, P, Y, eps, Pr=None):
    """predictMean and predictCovariance (GaussianProcessQ.ipp:101-186): with
    P_v = P P_r, mean P_r (sigma_e^2 Q + n_o P_v^T P_v)^-1 P_v^T sum_i y_i,
    covariance P_r sigma_e^2 (sigma_e^2 Q + n_o P_v^T P_v)^-1 P_r^T (P_r = I
    for integer beta)."""
    no = Y.shape[1]
    N = Q.shape[0]
    Pr = np.eye(N) if Pr is None else Pr
    Pv = P @ Pr
    Am = eps * eps * Q + no * Pv.T @ Pv
    mean = Pr @ np.linalg.solve(Am, Pv.T @ Y.sum(axis=1))
    cov = Pr @ (eps * eps * np.linalg.inv(Am)) @ Pr.T
    return mean, cov


# ------------------------------------------ first-order perturbation (ForwardPerturbation)
def fpert_update(rp, ci, va, F, coords, covvec, sigma_f, ell_f, P, ddofs, dblock, pcoords, sigma_k, ell_k,
                 reg_factor=1e-4):
    """ForwardPerturbation::updateSolution (statX/infer/ForwardPerturbation.hpp:108-190)
    with dense matrices: u = A^-1 F, R = P A^-1, the forcing part
    A^-1 C_f A^-T + u u^T with the full inverse (the reference fills the first
    n_s rows of its aInv, lines 127-141: equal when n_s = N), the material
    covariance C_k = k(x_e, x_d) of CovarianceMaterial (CovarianceMaterial.hpp:85-101)
    and the stiffness derivatives D_e = sum_ab dblock[e][a][b] e_{ddofs[e][a]} e_{ddofs[e][b]}^T:
      cov = R C_f R^T + sum_e sum_d C_k[e, d] R D_e (A^-1 C_f A^-T + u u^T) D_d^T R^T
            + (reg_factor sigma_k)^2 P P^T.
    Returns (P u, cov, u)."""
    lu = LU(rp, ci, va)
    u = lu.solve(F)
    N = len(u)
    R = lu.solve(np.ascontiguousarray(P.T), trans=True).T              # rows A^-T p_i
    Ainv = lu.solve(np.eye(N), trans=True).T                           # rows A^-T e_i = rows of A^-1
    force = Ainv @ covforcing_apply(coords, covvec, sigma_f, ell_f, np.ascontiguousarray(Ainv.T)) \
        + np.outer(u, u)
    pc = np.asarray(pcoords, np.float64).reshape(len(ddofs), -1)
    Ck = mismatch_cov(pc, sigma_k, ell_k, 0.0)
    ns = P.shape[0]
    RD, GE = [], []
    for e in range(len(ddofs)):
        De = np.zeros((N, N))
        De[np.ix_(ddofs[e], ddofs[e])] += dblock[e]
        RD.append(R @ De)
        GE.append(RD[-1] @ force)
    cov_m = np.zeros((ns, ns))
    for e in range(len(ddofs)):
        for d in range(len(ddofs)):
            cov_m += Ck[e, d] * GE[e] @ RD[d].T
    cov = covforcing_project(coords, covvec, sigma_f, ell_f, R) + cov_m + (reg_factor * sigma_k) ** 2 * P @ P.T
    return P @ u, cov, u
