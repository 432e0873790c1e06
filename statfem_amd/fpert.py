"""First-order perturbation of the forward solution under material (LHS) and
forcing (RHS) uncertainty on the GPU: ForwardPerturbation
(statX/infer/ForwardPerturbation.hpp:43-215) over sfh_fpert_* (include/statfem_hip.h)."""
import ctypes as C

import numpy as np

from . import hip
from . import mesh as M

# 1D Gauss-Legendre on [0, 1] (two points)
_G2 = (np.array([0.211324865405187, 0.788675134594813]), np.array([0.5, 0.5]))


def conductivity_derivatives(mesh, bc=None):
    """dA/dkappa_e for an element-wise conductivity of the Poisson operator:
    the unit-conductivity element stiffness on the element's nodes (LINE2,
    TRI3, QUAD4 with the 2 x 2 rule on [0,1]^2), rows of constrained DOFs zero
    (the row-zero form keeps those rows at the identity). Returns (dofs
    nel x nen, blocks nel x nen x nen) for setStiffnessDerivatives."""
    X = mesh.coords[mesh.conn]                                  # nel x nen x dim
    nel, nen = mesh.conn.shape
    if mesh.kind == M.LINE2:
        h = np.abs(X[:, 1, 0] - X[:, 0, 0])
        K = np.array([[1.0, -1.0], [-1.0, 1.0]])[None] / h[:, None, None]
    elif mesh.kind == M.TRI3:
        d1, d2 = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
        J = np.stack([d1, d2], axis=1)                          # rows: dx/dxi, dx/deta
        area = 0.5 * np.abs(np.linalg.det(J))
        gref = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])  # dN/d(xi, eta)
        G = np.einsum("eij,aj->eai", np.linalg.inv(J), gref)    # physical gradients
        K = area[:, None, None] * np.einsum("eai,ebi->eab", G, G)
    elif mesh.kind == M.QUAD4:
        K = np.zeros((nel, 4, 4))
        pts, wts = _G2
        for xi, wx in zip(pts, wts):
            for eta, wy in zip(pts, wts):
                # bilinear shape functions on [0,1]^2, nodes (0,0) (1,0) (1,1) (0,1)
                dN = np.array([[-(1 - eta), -(1 - xi)], [1 - eta, -xi], [eta, xi], [-eta, 1 - xi]])
                J = np.einsum("ai,ead->eid", dN, X)             # e x ref x phys
                G = np.einsum("ai,edi->ead", dN, np.linalg.inv(J))
                K += (wx * wy * np.abs(np.linalg.det(J)))[:, None, None] * np.einsum("ead,ebd->eab", G, G)
    else:
        raise ValueError("conductivity derivatives for LINE2, TRI3 and QUAD4 meshes")
    if bc is not None and len(bc):
        mask = np.zeros(mesh.nnodes, bool)
        mask[np.asarray(bc)] = True
        K = np.where(mask[mesh.conn][:, :, None], 0.0, K)
    return np.ascontiguousarray(mesh.conn, dtype=np.int32), np.ascontiguousarray(K)


class ForwardPerturbationHip:
    """system: assembled and constrained with the mean conductivity, its
    right-hand side the mean forcing; cols/vals: the projection P row by row;
    dofs/blocks: the stiffness derivatives; pcoords: the parameters'
    coordinates (element centroids) for CovarianceMaterial."""

    def __init__(self, system, cols, vals, dofs, blocks, pcoords, reg_factor=1e-4, nq1d=2):
        self.system = system
        self.ctx = system.ctx
        self.ns = len(cols)
        ns = self.ns
        pn = max(len(c) for c in cols)
        pc = np.zeros((ns, pn), np.int32)
        pv = np.zeros((ns, pn), np.float64)
        for i, (c, v) in enumerate(zip(cols, vals)):
            pc[i, :len(c)] = c
            pv[i, :len(c)] = v
            if len(c) < pn:
                pc[i, len(c):] = c[0] if len(c) else 0
        dofs = np.ascontiguousarray(dofs, np.int32)
        blocks = np.ascontiguousarray(blocks, np.float64)
        nk, nb = dofs.shape
        pcoords = np.ascontiguousarray(np.asarray(pcoords, np.float64).reshape(nk, -1))
        h = C.c_void_p()
        hip._check(hip.lib().sfh_fpert_create(system.h, nq1d, ns, pn, hip._p(pc), hip._p(pv), nk, nb, hip._p(dofs),
                                              hip._p(blocks), pcoords.shape[1], hip._p(pcoords), float(reg_factor),
                                              C.byref(h)))
        self.h = h
        self.ctx.adopt(self)
        self._mean = self._cov = None

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            hip.lib().sfh_fpert_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def updateSolution(self, params, opts=None):
        """params: {forcing.sigma, forcing.length, material.sigma, material.length}."""
        mean = np.zeros(self.ns)
        cov = np.zeros((self.ns, self.ns))
        st = hip.SolveStats()
        o = opts or hip.default_opts()
        hip._check(hip.lib().sfh_fpert_update(self.h, float(params["forcing.sigma"]), float(params["forcing.length"]),
                                              float(params["material.sigma"]), float(params["material.length"]),
                                              C.byref(o), hip._p(mean), hip._p(cov), None, C.byref(st)))
        self._mean, self._cov = mean, cov
        return mean, cov, st.as_dict()

    def giveProjectedMeanU(self):
        return self._mean.copy()

    def giveProjectedCovarU(self):
        return self._cov.copy()
