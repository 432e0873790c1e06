"""Python host mirror of the C ABI (include/statfem_hip.h) over ctypes.

Class and method names follow the reference interfaces they stand in for:
``SystemSolveHip`` ~ corlib::SystemSolveEigenSparse (+ MatrixAssembler over the
whole mesh), ``ForwardGaussRHSHip`` ~ statX::infer::ForwardGaussRHS,
``LikelihoodKOHHip`` ~ statX::infer::LikelihoodKOH (+ MismatchCovariance and
MCMC), ``gaussian_process`` ~ statX::infer::GaussianProcess.

The library must be built (statfem_amd/build.py); there is no CPU fallback:
a missing library or a failing device raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstatfem_hip.so")

SFH_LINE2, SFH_TRI3, SFH_QUAD4, SFH_HEX8 = 1, 2, 3, 4
POISSON, ELASTICITY, MATERN = 0, 1, 2
STATUS = {0: "SFH_OK", -1: "SFH_EARG", -2: "SFH_ENEGJAC", -3: "SFH_ESINGULAR", -4: "SFH_ENOTSPD",
          -5: "SFH_EQUADFORM", -6: "SFH_ENOCONV", -7: "SFH_EDEVICE", -8: "SFH_ENOMEM", -9: "SFH_ESTATE"}


class SfhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("max_inner", C.c_int), ("check_every", C.c_int), ("precision", C.c_int),
                ("rhs_tile", C.c_int), ("layout", C.c_int), ("precond", C.c_int), ("graph", C.c_int),
                ("lanes", C.c_int), ("mg_split", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_iters", C.c_int), ("max_relres", C.c_double),
                ("ms", C.c_double), ("converged", C.c_int), ("layout", C.c_int), ("precond", C.c_int),
                ("levels", C.c_int), ("graph_iters", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
v, i32, dbl, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
_SIGS = {
    "sfh_status_string": (C.c_char_p, [i32]),
    "sfh_last_error": (C.c_char_p, []),
    "sfh_version": (C.c_char_p, []),
    "sfh_ctx_create": (i32, [i32, C.POINTER(v)]),
    "sfh_ctx_destroy": (i32, [v]),
    "sfh_ctx_sync": (i32, [v]),
    "sfh_ctx_info": (i32, [v, C.POINTER(i32), C.POINTER(i32), C.c_char_p, i32]),
    "sfh_malloc": (i32, [v, sz, C.POINTER(v)]),
    "sfh_free": (i32, [v, v]),
    "sfh_memcpy_h2d": (i32, [v, v, v, sz]),
    "sfh_memcpy_d2h": (i32, [v, v, v, sz]),
    "sfh_memcpy_d2d": (i32, [v, v, v, sz]),
    "sfh_memset": (i32, [v, v, i32, sz]),
    "sfh_system_create_triplets": (i32, [v, i32, C.c_longlong, v, v, v, C.POINTER(v)]),
    "sfh_system_set_triplet_values": (i32, [v, C.c_longlong, v]),
    "sfh_system_set_rhs": (i32, [v, v]),
    "sfh_system_set_segment_grids": (i32, [v, i32, v, v, v]),
    "sfh_solve_batched": (i32, [v, i32, v, i32, i32, v, v, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "sfh_gram_segmented": (i32, [v, i32, v, i32, i32, v, i32, v, i32, v]),
    "sfh_covforcing_grid": (i32, [v, i32, i32, i32, v, v, v, v, dbl, dbl, i32, v, v]),
    "sfh_covforcing_batched": (i32, [v, i32, v, v, v, v, v, v, v, i32, v, v]),
    "sfh_dof_gather": (i32, [v, i32, i32, i32, i32, v, v, v]),
    "sfh_dof_scatter": (i32, [v, i32, i32, i32, i32, v, v, v]),
    "sfh_row_scale": (i32, [v, i32, i32, v, v, v]),
    "sfh_lik_logdet_batch": (i32, [v, i32, v, v, v]),
    "sfh_timer_start": (i32, [v]),
    "sfh_timer_stop": (i32, [v, C.POINTER(dbl)]),
    "sfh_system_create": (i32, [v, i32, i32, i32, v, i32, v, C.POINTER(v)]),
    "sfh_system_destroy": (i32, [v]),
    "sfh_system_info": (i32, [v, C.POINTER(i32), C.POINTER(C.c_longlong)]),
    "sfh_system_pattern": (i32, [v, v, v]),
    "sfh_system_set_grid": (i32, [v, i32, i32, i32, v, v, v]),
    "sfh_assemble_matrix": (i32, [v, i32, i32, v, v, dbl, dbl, dbl]),
    "sfh_assemble_rhs": (i32, [v, i32, v, v, dbl, i32]),
    "sfh_lumped_mass": (i32, [v, i32, v]),
    "sfh_apply_constraints": (i32, [v, i32, v, v, dbl, dbl]),
    "sfh_clear_matrix": (i32, [v]),
    "sfh_clear_rhs": (i32, [v]),
    "sfh_matrix_values": (i32, [v, v]),
    "sfh_rhs": (i32, [v, v]),
    "sfh_system_device_ptrs": (i32, [v, C.POINTER(v), C.POINTER(v), C.POINTER(v), C.POINTER(v)]),
    "sfh_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "sfh_solve": (i32, [v, i32, i32, v, v, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "sfh_solve_rhs": (i32, [v, v, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "sfh_kernel_sampling": (i32, [v, i32, i32, i32]),
    "sfh_kernel_times": (i32, [v, v, C.POINTER(C.c_longlong)]),
    "sfh_assembly_times": (i32, [v, v, C.POINTER(C.c_longlong)]),
    "sfh_covforcing_apply": (i32, [v, v, dbl, dbl, i32, v, v, i32]),
    "sfh_gram": (i32, [v, i32, i32, i32, v, i32, v, i32, v]),
    "sfh_forward_gauss": (i32, [v, i32, i32, v, v, dbl, dbl, dbl, C.POINTER(SolveOpts), v, v, v, v,
                                C.POINTER(SolveStats)]),
    "sfh_lik_create": (i32, [v, i32, i32, i32, v, v, v, v, dbl, C.POINTER(v)]),
    "sfh_lik_destroy": (i32, [v]),
    "sfh_lik_loglik_batch": (i32, [v, i32, v, v, v]),
    "sfh_lik_mismatch": (i32, [v, dbl, dbl, v]),
    "sfh_lik_post_proj": (i32, [v, dbl, dbl, dbl, v, v]),
    "sfh_lik_post_field": (i32, [v, dbl, dbl, dbl, i32, v, v, v]),
    "sfh_lik_post_field_cov": (i32, [v, dbl, dbl, dbl, i32, v, v, i32, v]),
    "sfh_system_set_covvec": (i32, [v, v]),
    "sfh_mcmc": (i32, [v, i32, i32, v, v, C.c_uint64, i32, v, v, v, v]),
    "sfh_lik_typical_readings": (i32, [v, dbl, dbl, dbl, i32, C.c_uint64, v]),
    "sfh_spmv": (i32, [v, i32, v, v]),
    "sfh_matern_tau": (i32, [dbl, dbl, dbl, i32, C.POINTER(dbl)]),
    "sfh_precision_apply": (i32, [v, v, dbl, dbl, dbl, i32, v, v]),
    "sfh_frac_apply": (i32, [v, v, i32, v, i32, i32, v, v]),
    "sfh_precision_apply_rational": (i32, [v, v, dbl, dbl, dbl, i32, v, i32, v, v]),
    "sfh_mlp_create": (i32, [v, C.c_uint64, C.POINTER(v)]),
    "sfh_mlp_destroy": (i32, [v]),
    "sfh_mlp_set_normalization": (i32, [v, v]),
    "sfh_mlp_params": (i32, [v, v, v, v]),
    "sfh_mlp_set_params": (i32, [v, v]),
    "sfh_mlp_grad": (i32, [v, i32, v, v, v, C.POINTER(dbl)]),
    "sfh_mlp_train": (i32, [v, i32, v, v, i32, dbl, C.POINTER(dbl)]),
    "sfh_mlp_predict": (i32, [v, i32, v, v]),
    "sfh_mcmc_da": (i32, [v, v, i32, i32, i32, v, v, C.c_uint64, i32, v, v, v, v, v]),
    "sfh_gp": (i32, [v, i32, i32, v, v, dbl, dbl, dbl, i32, v, v, v, v]),
    "sfh_gp_predict_cov": (i32, [v, i32, i32, v, v, dbl, dbl, dbl, i32, v, v]),
    "sfh_gp_noalloc": (i32, [v, i32, i32, v, v, dbl, dbl, dbl, i32, v, v, v, v, v]),
    "sfh_lhs_create": (i32, [v, i32, v, i32, i32, v, dbl, dbl, dbl, i32, i32, v, i32, i32, v, v, C.POINTER(v)]),
    "sfh_lhs_destroy": (i32, [v]),
    "sfh_lhs_set_data": (i32, [v, i32, v, i32, v, dbl, dbl, dbl, i32, v, v]),
    "sfh_lhs_update": (i32, [v, v, C.POINTER(SolveOpts), v, C.POINTER(SolveStats)]),
    "sfh_lhs_fix_mean": (i32, [v, i32]),
    "sfh_lhs_log_target": (i32, [v, i32, v, i32, C.POINTER(SolveOpts), v, v, C.POINTER(SolveStats)]),
    "sfh_lhs_mcmc": (i32, [v, i32, v, v, C.c_uint64, i32, C.POINTER(SolveOpts), v, v, C.POINTER(SolveStats)]),
    "sfh_gpq_create": (i32, [v, i32, i32, i32, v, v, i32, v, C.POINTER(v)]),
    "sfh_gpq_destroy": (i32, [v]),
    "sfh_gpq_set_rational": (i32, [v, dbl, i32, v, v]),
    "sfh_gpq_loglik": (i32, [v, i32, v, C.POINTER(SolveOpts), v, v, C.POINTER(SolveStats)]),
    "sfh_gpq_predict": (i32, [v, v, C.POINTER(SolveOpts), v, v, C.POINTER(dbl), C.POINTER(SolveStats)]),
    "sfh_fpert_create": (i32, [v, i32, i32, i32, v, v, i32, i32, v, v, i32, v, dbl, C.POINTER(v)]),
    "sfh_fpert_destroy": (i32, [v]),
    "sfh_fpert_update": (i32, [v, dbl, dbl, dbl, dbl, C.POINTER(SolveOpts), v, v, v, C.POINTER(SolveStats)]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SfhError(rc, lib().sfh_last_error().decode())


def _h(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_opts(**kw):
    o = SolveOpts()
    lib().sfh_solve_opts_default(C.byref(o))
    for k, val in kw.items():
        setattr(o, k, val)
    return o


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().sfh_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def adopt(self, child):
        self._children.add(child)

    def close(self):
        if getattr(self, "h", None):
            # handles that live on this context go first (their destroy uses the stream)
            for ch in list(self._children):
                ch.close()
            lib().sfh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        _check(lib().sfh_ctx_sync(self.h))

    def info(self):
        """{"device", "num_cu", "arch"} of the context's GPU"""
        d, n = C.c_int(), C.c_int()
        buf = C.create_string_buffer(64)
        _check(lib().sfh_ctx_info(self.h, C.byref(d), C.byref(n), buf, 64))
        return {"device": d.value, "num_cu": n.value, "arch": buf.value.decode()}

    def empty(self, shape, dtype=np.float64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a, dtype=np.float64):
        a = _h(a, dtype)
        d = DeviceArray(self, a.shape, dtype)
        d.upload(a)
        return d

    def kernel_sampling(self, every=16, max_samples=4096, min_rhs=0):
        """Sample the device time of the two inner-PCG kernels inside solves of
        at least min_rhs right-hand sides (per column tile)."""
        _check(lib().sfh_kernel_sampling(self.h, every, max_samples, min_rhs))

    def kernel_times(self):
        ms = np.zeros(3)
        n = C.c_longlong()
        _check(lib().sfh_kernel_times(self.h, _p(ms), C.byref(n)))
        return {"dir_spmv": float(ms[0]), "cg_update": float(ms[1]), "vcycle": float(ms[2]), "samples": n.value}

    def assembly_times(self):
        """Average device time (ms) of the assembly kernels per sfh_assemble_matrix
        call while kernel sampling is on, and the number of calls."""
        ms = np.zeros(1)
        n = C.c_longlong()
        _check(lib().sfh_assembly_times(self.h, _p(ms), C.byref(n)))
        return float(ms[0]), n.value

    def timer_start(self):
        _check(lib().sfh_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_double()
        _check(lib().sfh_timer_stop(self.h, C.byref(ms)))
        return ms.value


class DeviceArray:
    """Device buffer from sfh_malloc; freed with the object."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().sfh_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            lib().sfh_free(self.ctx.h, self.ptr)
            self.ptr = None

    def upload(self, a):
        a = _h(a, self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().sfh_memcpy_h2d(self.ctx.h, self.ptr, _p(a), self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().sfh_memcpy_d2h(self.ctx.h, _p(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        _check(lib().sfh_memset(self.ctx.h, self.ptr, 0, self.nbytes))


class SystemSolveHip:
    """GPU FE system: mesh, CSR pattern, assembled operator, RHS and constraints.

    Mirrors SystemSolveEigenSparse (corlib/SystemSolveEigenSparse.hpp:54-122):
    ``applyConstraints``, ``solveSystem``, ``solveSystemTranspose``,
    ``giveSolution``-style host copies; whole-mesh assembly replaces the
    per-element MatrixAssembler callback (corlib/MatrixAssembler.hpp:198-232).
    """

    def __init__(self, ctx, mesh, dofpn=1):
        self.ctx = ctx
        self.mesh = mesh
        self.dofpn = dofpn
        h = C.c_void_p()
        coords = _h(mesh.coords)
        conn = _h(mesh.conn, np.int32)
        _check(lib().sfh_system_create(ctx.h, mesh.kind, dofpn, mesh.nnodes, _p(coords), mesh.nel,
                                       _p(conn), C.byref(h)))
        self.h = h
        ctx.adopt(self)
        n = C.c_int()
        nnz = C.c_longlong()
        _check(lib().sfh_system_info(h, C.byref(n), C.byref(nnz)))
        self.ndof, self.nnz = n.value, nnz.value
        self._keep = []

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().sfh_system_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def set_grid(self, grid=None):
        g = grid if grid is not None else self.mesh.grid
        axes = [_h(a) for a in g] + [None] * (3 - len(g))
        ns = [len(a) if a is not None else 1 for a in axes]
        _check(lib().sfh_system_set_grid(self.h, ns[0], ns[1], ns[2], _p(axes[0]), _p(axes[1]), _p(axes[2])))

    def pattern(self):
        rp = np.zeros(self.ndof + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        _check(lib().sfh_system_pattern(self.h, _p(rp), _p(ci)))
        return rp, ci

    def assemble(self, physics, p1=1.0, p2=0.0, nq1d=2, factor=1.0):
        """p1/p2: scalars or per-element arrays (host or DeviceArray)."""
        args = []
        scal = []
        for p in (p1, p2):
            if isinstance(p, DeviceArray):
                args.append(p.ptr); scal.append(0.0)
            elif np.ndim(p) == 0:
                args.append(None); scal.append(float(p))
            else:
                d = self.ctx.to_device(p)
                self._keep.append(d)
                args.append(d.ptr); scal.append(0.0)
        _check(lib().sfh_assemble_matrix(self.h, physics, nq1d, args[0], args[1], scal[0], scal[1], factor))
        self._keep = []

    def assemble_rhs(self, nq1d=2, fconst=None, fnodal=None, factor=1.0, accumulate=False):
        fc = None if fconst is None else _h(np.broadcast_to(fconst, (self.dofpn,)))
        fn = None
        if fnodal is not None:
            fn = fnodal if isinstance(fnodal, DeviceArray) else self.ctx.to_device(fnodal)
        _check(lib().sfh_assemble_rhs(self.h, nq1d, _p(fc), None if fn is None else fn.ptr, factor,
                                      int(accumulate)))

    def lumped_mass(self, nq1d=2):
        out = self.ctx.empty((self.ndof,))
        _check(lib().sfh_lumped_mass(self.h, nq1d, out.ptr))
        return out

    def applyConstraints(self, dofs, values=None, factor=1.0, scalar=1.0):
        d = _h(dofs, np.int32)
        vals = None if values is None else _h(values)
        _check(lib().sfh_apply_constraints(self.h, len(d), _p(d), _p(vals), factor, scalar))

    def values(self):
        out = np.zeros(self.nnz)
        _check(lib().sfh_matrix_values(self.h, _p(out)))
        return out

    def rhs(self):
        out = np.zeros(self.ndof)
        _check(lib().sfh_rhs(self.h, _p(out)))
        return out

    def solve(self, B, trans=False, opts=None):
        """B: (N, k) host array or DeviceArray. Returns (X DeviceArray, stats dict)."""
        Bd = B if isinstance(B, DeviceArray) else self.ctx.to_device(np.asarray(B).reshape(self.ndof, -1))
        k = int(np.prod(Bd.shape) // self.ndof)
        X = self.ctx.empty((self.ndof, k))
        st = SolveStats()
        o = opts or default_opts()
        _check(lib().sfh_solve(self.h, int(trans), k, Bd.ptr, X.ptr, C.byref(o), C.byref(st)))
        return X, st.as_dict()

    def solveSystem(self, opts=None):
        X = self.ctx.empty((self.ndof,))
        st = SolveStats()
        o = opts or default_opts()
        _check(lib().sfh_solve_rhs(self.h, X.ptr, C.byref(o), C.byref(st)))
        return X, st.as_dict()

    def solveSystemTranspose(self, rhs, opts=None):
        X, st = self.solve(np.asarray(rhs).reshape(-1, 1), trans=True, opts=opts)
        return X.download().ravel(), st

    def spmv(self, X):
        Xd = X if isinstance(X, DeviceArray) else self.ctx.to_device(np.asarray(X).reshape(self.ndof, -1))
        k = int(np.prod(Xd.shape) // self.ndof)
        Y = self.ctx.empty((self.ndof, k))
        _check(lib().sfh_spmv(self.h, k, Xd.ptr, Y.ptr))
        return Y

    def precision_apply(self, mass, sigma, kappa, nu, X, bot=None):
        """PrecisionMatrixRational::givePrecisionMatrix applied to X: integer
        beta, or a fractional one with its rational denominator factors bot
        (b, d_0 ..): Q_v X = P_l^T Q P_l X."""
        Xd = X if isinstance(X, DeviceArray) else self.ctx.to_device(np.asarray(X).reshape(self.ndof, -1))
        k = int(np.prod(Xd.shape) // self.ndof)
        md = mass if isinstance(mass, DeviceArray) else self.ctx.to_device(mass)
        Y = self.ctx.empty((self.ndof, k))
        if bot is None:
            _check(lib().sfh_precision_apply(self.h, md.ptr, sigma, kappa, nu, k, Xd.ptr, Y.ptr))
        else:
            b = _h(np.asarray(bot, np.float64))
            _check(lib().sfh_precision_apply_rational(self.h, md.ptr, sigma, kappa, nu, len(b) - 1, _p(b), k,
                                                      Xd.ptr, Y.ptr))
        return Y

    def frac_apply(self, mass, roots, X, transpose=False):
        """P_l or P_r (giveLeftFracMatrix / giveRightFracMatrix) applied to X:
        roots[0] prod_j (I - roots[j] M^-1 A) X, or the transpose."""
        Xd = X if isinstance(X, DeviceArray) else self.ctx.to_device(np.asarray(X).reshape(self.ndof, -1))
        k = int(np.prod(Xd.shape) // self.ndof)
        md = mass if isinstance(mass, DeviceArray) else self.ctx.to_device(mass)
        r = _h(np.asarray(roots, np.float64))
        Y = self.ctx.empty((self.ndof, k))
        _check(lib().sfh_frac_apply(self.h, md.ptr, len(r) - 1, _p(r), int(bool(transpose)), k, Xd.ptr, Y.ptr))
        return Y

    def covforcing_apply(self, covvec, sigma, ell, W, mode=0):
        Wd = W if isinstance(W, DeviceArray) else self.ctx.to_device(np.asarray(W).reshape(self.ndof, -1))
        k = int(np.prod(Wd.shape) // self.ndof)
        cv = covvec if isinstance(covvec, DeviceArray) else self.ctx.to_device(covvec)
        V = self.ctx.empty((self.ndof, k))
        _check(lib().sfh_covforcing_apply(self.h, cv.ptr, sigma, ell, k, Wd.ptr, V.ptr, mode))
        return V


def solve_probe(system, B, X, k, trans=False, opts=None, seg=None):
    """Run a (possibly batched) solve that may stop at its iteration cap and
    return its statistics (for timing probes): SFH_ENOCONV is not an error here."""
    st = SolveStats()
    o = opts or default_opts()
    if seg is None:
        rc = lib().sfh_solve(system.h, int(trans), k, B.ptr, X.ptr, C.byref(o), C.byref(st))
    else:
        seg = _h(seg, np.int32)
        rc = lib().sfh_solve_batched(system.h, len(seg) - 1, _p(seg), int(trans), k, B.ptr, X.ptr, C.byref(o),
                                     C.byref(st))
    if rc not in (0, -6):
        _check(rc)
    return st.as_dict()


def dof_gather(ctx, w, src, stride, comp, out=None):
    """out[j, c] = w[j] src[j*stride + comp, c] (device arrays)."""
    n = w.shape[0]
    k = int(np.prod(src.shape) // (n * stride))
    out = out if out is not None else ctx.empty((n, k))
    _check(lib().sfh_dof_gather(ctx.h, n, k, stride, comp, w.ptr, src.ptr, out.ptr))
    return out


def dof_scatter(ctx, w, src, stride, comp, out=None):
    """out[j*stride + q, c] = (q == comp) w[j] src[j, c] (device arrays)."""
    n = w.shape[0]
    k = int(np.prod(src.shape) // n)
    out = out if out is not None else ctx.empty((n * stride, k))
    _check(lib().sfh_dof_scatter(ctx.h, n, k, stride, comp, w.ptr, src.ptr, out.ptr))
    return out


def row_scale(ctx, w, X, out=None):
    """out = diag(w) X (device arrays; out may be X)."""
    n = w.shape[0]
    k = int(np.prod(X.shape) // n)
    out = out if out is not None else ctx.empty(X.shape)
    _check(lib().sfh_row_scale(ctx.h, n, k, w.ptr, X.ptr, out.ptr))
    return out


def gram(ctx, X, V):
    N, m = X.shape
    n = V.shape[1]
    Cd = ctx.empty((m, n))
    _check(lib().sfh_gram(ctx.h, N, m, n, X.ptr, m, V.ptr, n, Cd.ptr))
    return Cd.download()


class ForwardGaussRHSHip:
    """ForwardGaussRHS::updateSolution (statX/infer/ForwardGaussRHS.hpp:95-139)
    with the projection P given as sparse rows (cols, vals) of shape (ns, pn)."""

    def __init__(self, system, cols, vals, reg_eps=1e-4):
        self.sys = system
        self.cols = _h(cols, np.int32)
        self.vals = _h(vals)
        self.ns, self.pn = self.cols.shape
        self.reg_eps = reg_eps
        self.field = None
        self.ubar = None

    def updateSolution(self, sigma_f, ell_f, want_field=False, opts=None):
        ns = self.ns
        mean = np.zeros(ns)
        Cu = np.zeros((ns, ns))
        if want_field and getattr(self, "field", None) is None:   # kept across updates
            self.field = self.sys.ctx.empty((self.sys.ndof, ns))
            self.ubar = self.sys.ctx.empty((self.sys.ndof,))
        st = SolveStats()
        o = opts or default_opts()
        _check(lib().sfh_forward_gauss(self.sys.h, ns, self.pn, _p(self.cols), _p(self.vals), sigma_f, ell_f,
                                       self.reg_eps, C.byref(o), _p(mean), _p(Cu),
                                       self.field.ptr if want_field else None,
                                       self.ubar.ptr if want_field else None, C.byref(st)))
        self.mean, self.Cu = mean, Cu
        return mean, Cu, st.as_dict()


class LikelihoodKOHHip:
    """LikelihoodKOH + MismatchCovariance (statX/infer/LikelihoodKOH.ipp,
    statX/infer/MismatchCovariance.hpp). Y: (no, ns) snapshot-major."""

    def __init__(self, ctx, xs, Cu, mean, Y, eps):
        xs = _h(np.asarray(xs, dtype=np.float64).reshape(len(mean), -1))
        Y = _h(np.atleast_2d(Y))
        self.ctx = ctx
        self.ns = len(mean)
        self.no = Y.shape[0]
        self.dim = xs.shape[1]
        h = C.c_void_p()
        _check(lib().sfh_lik_create(ctx.h, self.ns, self.no, self.dim, _p(xs), _p(_h(Cu)), _p(_h(mean)), _p(Y),
                                    eps, C.byref(h)))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().sfh_lik_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def giveLogLikelihood(self, theta):
        """theta: (M, 3) natural-scale (rho, mismatch.sigma, mismatch.length)."""
        th = _h(np.atleast_2d(theta))
        M = th.shape[0]
        out = np.zeros(M)
        st = np.zeros(M, np.int32)
        _check(lib().sfh_lik_loglik_batch(self.h, M, _p(th), _p(out), _p(st)))
        return out, st

    def giveLogDeterminant(self, theta):
        """log det(K_s + rho^2 Cu) for (M, 3) natural-scale hyperparameters."""
        th = _h(np.atleast_2d(theta))
        M = th.shape[0]
        out = np.zeros(M)
        st = np.zeros(M, np.int32)
        _check(lib().sfh_lik_logdet_batch(self.h, M, _p(th), _p(out), _p(st)))
        return out, st

    def mismatch(self, sd, ld):
        Ks = np.zeros((self.ns, self.ns))
        _check(lib().sfh_lik_mismatch(self.h, sd, ld, _p(Ks)))
        return Ks

    def giveProjFEconditionedData(self, rho, sd, ld):
        m = np.zeros(self.ns)
        c = np.zeros((self.ns, self.ns))
        _check(lib().sfh_lik_post_proj(self.h, rho, sd, ld, _p(m), _p(c)))
        return m, c

    def giveFEconditionedDataWoodbury(self, rho, sd, ld, field, ubar):
        N = field.shape[0]
        out = self.ctx.empty((N,))
        _check(lib().sfh_lik_post_field(self.h, rho, sd, ld, N, field.ptr, ubar.ptr, out.ptr))
        return out

    def giveFEconditionedCovarianceWoodbury(self, rho, sd, ld, field, prior, full=True):
        """Posterior field covariance C_u - (C_u P^T) S^-1 (P C_u) of
        giveFEconditionedDataWoodbury (statX/infer/LikelihoodKOH.ipp:459-488)
        from field = C_u P^T (device N x ns) and the prior: the N x N prior
        covariance (full) or its diagonal (full=False, returns the variance)."""
        N = field.shape[0]
        pr = prior if isinstance(prior, DeviceArray) else self.ctx.to_device(prior)
        out = self.ctx.empty((N, N) if full else (N,))
        _check(lib().sfh_lik_post_field_cov(self.h, rho, sd, ld, N, field.ptr, pr.ptr, int(bool(full)), out.ptr))
        return out

    def giveTypicalReadings(self, rho, sd, ld, nr, seed=0):
        Y = np.zeros((nr, self.ns))
        _check(lib().sfh_lik_typical_readings(self.h, rho, sd, ld, nr, seed, _p(Y)))
        return Y

    def mcmc(self, theta0, propsd, steps, seed=0, prior_mode=0, prior_mu=(0, 0, 0), prior_var=(1, 1, 1),
             keep_samples=True):
        th0 = _h(np.atleast_2d(theta0))
        chains = th0.shape[0]
        samples = np.zeros((chains, steps, 4)) if keep_samples else None
        acc = np.zeros(chains)
        _check(lib().sfh_mcmc(self.h, chains, steps, _p(th0), _p(_h(propsd)), seed, prior_mode,
                              _p(_h(prior_mu)), _p(_h(prior_var)), _p(samples), _p(acc)))
        return samples, acc


def gaussian_process(ctx, X, y, sigma, ell, eps, Xp):
    y = _h(y)
    X = _h(X).reshape(len(y), -1)
    dim = X.shape[1]
    Xp = _h(Xp).reshape(-1, dim)
    logl = C.c_double()
    mean = np.zeros(Xp.shape[0])
    var = np.zeros(Xp.shape[0])
    _check(lib().sfh_gp(ctx.h, len(y), dim, _p(X), _p(y), sigma, ell, eps, Xp.shape[0], _p(Xp), C.byref(logl),
                        _p(mean), _p(var)))
    return logl.value, mean, var


def gp_predict_covariance(ctx, X, y, sigma, ell, eps, Xp):
    """GaussianProcess::predictCovariance (statX/infer/GaussianProcess.hpp:160-190)
    at the test locations Xp (np x dim); returns the np x np covariance."""
    y = _h(y)
    X = _h(X).reshape(len(y), -1)
    dim = X.shape[1]
    Xp = _h(Xp).reshape(-1, dim)
    cov = np.zeros((Xp.shape[0], Xp.shape[0]))
    _check(lib().sfh_gp_predict_cov(ctx.h, len(y), dim, _p(X), _p(y), sigma, ell, eps, Xp.shape[0], _p(Xp),
                                    _p(cov)))
    return cov


def matern_tau(sigma, kappa, nu, dim):
    t = C.c_double()
    _check(lib().sfh_matern_tau(sigma, kappa, nu, dim, C.byref(t)))
    return t.value


class SurrogateMLP:
    """MLP 3 -> 64 -> 64 -> 1 (tanh) surrogate of the KOH log-likelihood;
    BF16 weights, FP32 accumulation, trained on the GPU."""
    NP = 4481

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().sfh_mlp_create(ctx.h, seed, C.byref(h)))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().sfh_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def params(self):
        m = np.zeros(self.NP, np.float32)
        b = np.zeros(self.NP, np.uint16)
        nrm = np.zeros(8)
        _check(lib().sfh_mlp_params(self.h, _p(m), _p(b), _p(nrm)))
        return m, b, nrm

    def set_params(self, master):
        _check(lib().sfh_mlp_set_params(self.h, _p(_h(master, np.float32))))

    def set_normalization(self, norm8):
        _check(lib().sfh_mlp_set_normalization(self.h, _p(_h(norm8))))

    def grad(self, theta, logl):
        th = _h(np.atleast_2d(theta))
        g = np.zeros(self.NP, np.float32)
        loss = C.c_double()
        _check(lib().sfh_mlp_grad(self.h, th.shape[0], _p(th), _p(_h(logl)), _p(g), C.byref(loss)))
        return g, loss.value

    def train(self, theta, logl, steps=2000, lr=3e-3):
        th = _h(np.atleast_2d(theta))
        loss = C.c_double()
        _check(lib().sfh_mlp_train(self.h, th.shape[0], _p(th), _p(_h(logl)), steps, lr, C.byref(loss)))
        return loss.value

    def predict(self, theta):
        th = _h(np.atleast_2d(theta))
        out = np.zeros(th.shape[0])
        _check(lib().sfh_mlp_predict(self.h, th.shape[0], _p(th), _p(out)))
        return out


def mcmc_da(lik, mlp, theta0, propsd, steps, sub, seed=0, prior_mode=0, prior_mu=(0, 0, 0), prior_var=(1, 1, 1),
            keep_samples=True):
    th0 = _h(np.atleast_2d(theta0))
    chains = th0.shape[0]
    samples = np.zeros((chains, steps, 4)) if keep_samples else None
    acc = np.zeros(chains)
    sacc = np.zeros(chains)
    _check(lib().sfh_mcmc_da(lik.h, mlp.h, chains, steps, sub, _p(th0), _p(_h(propsd)), seed, prior_mode,
                             _p(_h(prior_mu)), _p(_h(prior_var)), _p(samples), _p(acc), _p(sacc)))
    return samples, acc, sacc
