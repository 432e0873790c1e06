"""Pins the CPU FP64 oracle against the reference's golden vectors, published
known answers and closed forms, so that it can serve as the parity checker of
the HIP path. Each test names what it pins."""
import os

import numpy as np
import pytest

from statfem_amd import mesh as M

from .conftest import GOLDEN


# ------------------------------------------------------------------ RNG
def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32-10 (Salmon et al., SC'11)
    assert [hex(v) for v in orc.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(v) for v in orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(v) for v in orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                       [0xA4093822, 0x299F31D0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_mcmc_draws_are_standard_normal(orc):
    z = np.array([orc.mcmc_draws(7, c, t)[0] for c in range(4) for t in range(2500)]).ravel()
    u = np.array([orc.mcmc_draws(7, c, t)[1] for c in range(4) for t in range(2500)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    assert abs(u.mean() - 0.5) < 0.02 and u.min() > 0 and u.max() < 1


# ------------------------------------------------------------ quadrature
def test_quadrature_constants_and_order(orc):
    w, xi = orc.quad_rule(M.QUAD4, 2)
    # corlib/GaussLegendre.hpp:64-66 constants; TensorQuadrature: last axis outermost
    assert np.allclose(w, 0.25, atol=0) and xi[0, 0] == 0.788675134594813
    assert xi[1, 0] == 0.211324865405187 and xi[1, 1] == 0.788675134594813
    w3, _ = orc.quad_rule(M.LINE2, 3)
    assert w3[0] == 0.277777777777777 and w3[1] == 0.444444444444444
    wt, xt = orc.quad_rule(M.TRI3, 1)
    assert wt[0] == 0.5 and xt[0, 0] == 0.333333333333333
    wh, _ = orc.quad_rule(M.HEX8, 2)
    assert len(wh) == 8 and abs(wh.sum() - 1.0) < 1e-15


# ------------------------------------------------------- closed forms
def test_q1_uniform_grid_nine_point_stencil(orc):
    m = M.grid2d(6)
    rp, ci, va, _ = orc.assemble(m, orc.POISSON, 1.0)
    i = 3 * 7 + 3   # interior node
    row = dict(zip(ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]]))
    assert len(row) == 9
    assert abs(row[i] - 8.0 / 3.0) < 1e-14
    for j, v in row.items():
        if j != i:
            assert abs(v + 1.0 / 3.0) < 1e-14
    # row sums vanish before constraints; symmetry
    A = np.zeros((m.nnodes, m.nnodes))
    for r in range(m.nnodes):
        A[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    assert np.abs(A.sum(1)).max() < 1e-13
    assert np.abs(A - A.T).max() == 0.0


def test_p1_right_triangles_five_point_stencil_with_explicit_zeros(orc):
    m = M.grid2d(5, kind=M.TRI3)
    rp, ci, va, _ = orc.assemble(m, orc.POISSON, 2.0)
    i = 2 * 6 + 2
    row = dict(zip(ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]]))
    # pattern holds the two diagonal couplings along the split direction as explicit zeros
    assert sorted(row) == sorted([i - 7, i - 6, i - 1, i, i + 1, i + 6, i + 7])
    assert abs(row[i] - 8.0) < 1e-13
    for j in (i - 6, i - 1, i + 1, i + 6):
        assert abs(row[j] + 2.0) < 1e-13
    assert abs(row[i - 7]) < 1e-15 and abs(row[i + 7]) < 1e-15


def test_shape_integrals_sum_to_measure(orc):
    for mesh, nq in ((M.grid2d(7), 2), (M.grid2d(7, kind=M.TRI3), 1), (M.grid3d(3), 2), (M.line1d(9), 3)):
        F = orc.body_force(mesh, nq1d=nq, fconst=1.0)
        assert abs(F.sum() - 1.0) < 1e-13


def test_hex8_elasticity_rigid_body_modes_and_symmetry(orc):
    m = M.grid3d(2)
    rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, 1.0, 0.3)
    n = 3 * m.nnodes
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    assert np.abs(A - A.T).max() < 1e-15
    x = m.coords
    modes = []
    for c in range(3):
        u = np.zeros((m.nnodes, 3)); u[:, c] = 1; modes.append(u.ravel())
    for a, b in ((0, 1), (1, 2), (0, 2)):
        u = np.zeros((m.nnodes, 3)); u[:, a] = -x[:, b]; u[:, b] = x[:, a]; modes.append(u.ravel())
    for u in modes:
        assert np.abs(A @ u).max() < 1e-13
    # a pure stretch is not in the kernel, and each DOF row spans 27 nodes x 3
    u = np.zeros((m.nnodes, 3)); u[:, 0] = x[:, 0]
    assert np.abs(A @ u.ravel()).max() > 1e-3
    centre = (1 * 3 + 1) * 3 + 1
    assert rp[3 * centre + 1] - rp[3 * centre] == 81


def _dense(rp, ci, va, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    return A


def _hex8_meshes():
    """a unit cube, a 2^3 block, and a 3 x 2 x 2 block on the box [0,2]x[0,1]x[0,0.5]
    with its interior nodes moved (general trilinear geometry)"""
    rng = np.random.default_rng(11)
    box = M.grid3d(3, 2, 2)
    X = box.coords * np.array([2.0, 1.0, 0.5])
    inner = np.all((box.coords > 1e-9) & (box.coords < 1 - 1e-9), axis=1)
    X[inner] += rng.uniform(-0.08, 0.08, size=X[inner].shape)
    return [(M.grid3d(1), 1.0), (M.grid3d(2), 1.0), (M.Mesh(M.HEX8, X, box.conn), 1.0)]


def test_hex8_elasticity_strain_energy_closed_forms(orc):
    """Linear displacement fields u = G x are reproduced exactly by the
    isoparametric hex8 element, so 1/2 u^T K u equals the continuum strain
    energy 1/2 |Omega| (lambda tr(eps)^2 + 2 mu eps:eps), eps = sym(G), with the
    Lame constants of solid/material/svenant/SVenant.ipp:32-37. Uniaxial
    strain gives 1/2 (lambda + 2 mu) e^2 |Omega|, simple shear 1/2 mu g^2 |Omega|,
    dilatation 3/2 (3 lambda + 2 mu) e^2 |Omega|: a lambda/mu swap or a wrong
    Lame formula changes each of them."""
    E, nu = 2.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    e, g = 1e-2, 3e-2
    cases = [
        (np.diag([e, 0.0, 0.0]), 0.5 * (lam + 2 * mu) * e * e),
        (np.array([[0.0, g, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]), 0.5 * mu * g * g),
        (np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, g, 0.0]]), 0.5 * mu * g * g),
        (e * np.eye(3), 1.5 * (3 * lam + 2 * mu) * e * e),
    ]
    for m, _ in _hex8_meshes():
        vol = orc.body_force(m, nq1d=2, fconst=1.0).sum()
        n = 3 * m.nnodes
        rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, E, nu)
        A = _dense(rp, ci, va, n)
        for G, dens in cases:
            u = (m.coords @ G.T).ravel()
            energy = 0.5 * u @ A @ u
            assert abs(energy - dens * vol) <= 1e-12 * dens * vol, (energy, dens * vol)


def test_hex8_elasticity_patch_test(orc):
    """Constant strain is an equilibrium state: for u = G x (any G) the nodal
    forces K u vanish at every interior node, and the boundary forces equal the
    tractions sigma n of the constant stress sigma = lambda tr(eps) I + 2 mu eps
    integrated over the boundary (their sum is zero and their moment balances)."""
    E, nu = 1.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    rng = np.random.default_rng(5)
    G = rng.normal(size=(3, 3))
    eps = 0.5 * (G + G.T)
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    m = M.grid3d(3)
    rp, ci, va, _ = orc.assemble(m, orc.ELASTICITY, E, nu)
    A = _dense(rp, ci, va, 3 * m.nnodes)
    f = (A @ (m.coords @ G.T).ravel()).reshape(-1, 3)
    inner = np.all((m.coords > 1e-9) & (m.coords < 1 - 1e-9), axis=1)
    assert np.abs(f[inner]).max() < 1e-13 * np.abs(f).max()
    # on the face x = 1 the traction is sigma e_x; its resultant over the face is sigma e_x (area 1)
    face = np.abs(m.coords[:, 0] - 1.0) < 1e-12
    edge_free = face & np.all((m.coords[:, 1:] > 1e-9) & (m.coords[:, 1:] < 1 - 1e-9), axis=1)
    # an interior face node carries the traction times its tributary area h^2 = 1/9
    assert np.abs(f[edge_free] - sig[:, 0] / 9.0).max() < 1e-13
    assert np.abs(f.sum(axis=0)).max() < 1e-13


def test_matern_row_sums_are_kappa2_lumped_mass(orc):
    m = M.grid2d(5)
    kap = 3.0
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0)
    ml = orc.body_force(m, fconst=1.0)
    rows = np.add.reduceat(va, rp[:-1])
    assert np.abs(rows - kap * kap * ml).max() < 1e-13


# ------------------------------------------------------------ pattern
def test_pattern_matches_brute_force_union(orc):
    for mesh, dofpn in ((M.grid2d(4, 3, kind=M.TRI3), 1), (M.grid3d(2, 1, 2), 3), (M.line1d(5), 1)):
        phys = orc.ELASTICITY if dofpn == 3 else orc.POISSON
        rp, ci, _, _ = orc.assemble(mesh, phys, 1.0, 0.3)
        ed = mesh.element_dofs(dofpn)
        n = mesh.nnodes * dofpn
        S = [set() for _ in range(n)]
        for e in ed:
            for a in e:
                S[a].update(e.tolist())
        exp_rp = np.concatenate([[0], np.cumsum([len(s) for s in S])])
        assert np.array_equal(rp, exp_rp)
        assert np.array_equal(ci, np.concatenate([sorted(s) for s in S]))


def test_row_zero_dirichlet_keeps_pattern(orc):
    m = M.grid2d(4)
    bc = M.dirichlet_all_boundary(m)
    rp0, ci0, _, _ = orc.assemble(m, orc.POISSON, 1.0)
    F = np.ones(m.nnodes)
    rp, ci, va, F2 = orc.assemble(m, orc.POISSON, 1.0, bc_dofs=bc, bc_vals=np.full(len(bc), 0.5), F=F)
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    for c in bc:
        seg = slice(rp[c], rp[c + 1])
        assert np.all(va[seg][ci[seg] != c] == 0.0) and va[seg][ci[seg] == c][0] == 1.0
        assert F2[c] == 0.5


# --------------------------------------------------------------- solves
def test_direct_solves_match_dense(orc):
    m = M.grid2d(6, 5, kind=M.TRI3)
    bc = M.dirichlet_all_boundary(m)
    rng = np.random.default_rng(0)
    kap = np.exp(rng.normal(size=m.nel) * 0.3)
    rp, ci, va, _ = orc.assemble(m, orc.POISSON, kap, bc_dofs=bc)
    n = m.nnodes
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    B = rng.normal(size=(n, 3))
    lu = orc.LU(rp, ci, va)
    assert np.abs(lu.solve(B) - np.linalg.solve(A, B)).max() < 1e-12
    assert np.abs(lu.solve(B, trans=True) - np.linalg.solve(A.T, B)).max() < 1e-12
    mask = np.zeros(n, np.uint8); mask[bc] = 1
    ch = orc.Chol(rp, ci, va, mask)
    Bh = B.copy(); Bh[bc] = 0
    assert np.abs(ch.solve(Bh) - np.linalg.solve(A, Bh)).max() < 1e-12


# --------------------------------------------------------- GP golden
def _gp_fixture():
    g = os.path.join(GOLDEN, "gp")
    hp = dict(l.split() for l in open(os.path.join(g, "gpHps.dat")) if l.strip())
    X = np.loadtxt(os.path.join(g, "gpCoordinates.dat"), skiprows=1)
    y = np.array([float(hp[f"gaussianProcess.w{i}"]) for i in range(6)])
    Xp = np.loadtxt(os.path.join(g, "predictionCoord.dat"))
    mean = np.loadtxt(os.path.join(g, "watchDog.0.mk.predictionMean.dat"))
    var = np.loadtxt(os.path.join(g, "watchDog.0.mk.predictionVar.dat"))
    return hp, X, y, Xp, mean, var


def test_gp_golden(orc):
    hp, X, y, Xp, mean, var = _gp_fixture()
    ll, mu, v = orc.gp(X, y, float(hp["gaussianProcess.sigma"]), float(hp["gaussianProcess.length"]),
                       float(hp["gaussianProcess.eps"]), Xp)
    assert abs(ll - (-4.837171259)) < 1e-9          # watchDog.0.mk:19
    assert np.abs(mu - mean).max() < 1e-8            # watchDog.0.mk:18 TOL
    assert np.abs(v - var).max() < 1e-8


# ------------------------------------------- 1D statFEM end-to-end golden
def laplace_rhs_infer_problem(orc):
    """del2/apps/reference/laplaceRHSinfer: LINE2 bar, 3-point rule, f = pi^2/5,
    u(0) = u(1) = 0, forcing (0.3, 0.25), reg 1e-3 (inference.hpp:136)."""
    g = os.path.join(GOLDEN, "laplaceRHSinfer")
    m = M.read_smf(os.path.join(g, "bar.smf"))
    bc = M.fuzzy_boundary(m.coords, [0])
    F = orc.body_force(m, nq1d=3, fconst=np.pi ** 2 / 5)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, 1.0, nq1d=3, bc_dofs=bc,
                                 bc_vals=np.zeros(len(bc)), F=F)
    covvec = orc.body_force(m, nq1d=3, fconst=1.0)
    covvec[bc] = 0.0
    trip = np.loadtxt(os.path.join(g, "phiMat.dat"))
    ns = int(trip[:, 0].max()) + 1
    P = np.zeros((ns, m.nnodes))
    P[trip[:, 0].astype(int), trip[:, 1].astype(int)] = trip[:, 2]
    Y = np.loadtxt(os.path.join(g, "sensorValues.dat"), skiprows=1).T.copy()
    xs = np.loadtxt(os.path.join(g, "sensorCoordinates.dat"))[:, None]
    ref = np.loadtxt(os.path.join(g, "watchDog.0.mk.inferredFEcondData.dat"))
    return dict(mesh=m, bc=bc, rp=rp, ci=ci, va=va, F=F, covvec=covvec, P=P, Y=Y, xs=xs, ref=ref,
                sf=0.3, lf=0.25, reg=1e-3, eps=0.005,
                hp=(0.804412, 0.0191724, 0.0740285))


def test_laplace_rhs_infer_woodbury_golden(orc):
    p = laplace_rhs_infer_problem(orc)
    m = p["mesh"]
    lu = orc.LU(p["rp"], p["ci"], p["va"])
    Rfull = lu.solve(np.eye(m.nnodes), trans=True).T
    CuFull = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], Rfull) \
        + (p["sf"] * p["reg"]) ** 2 * np.eye(m.nnodes)
    ubar = Rfull @ p["F"]
    rho, sd, ld = p["hp"]
    Ks = orc.mismatch_cov(p["xs"], sd, ld, p["eps"])
    mean, cov = orc.post_woodbury(CuFull, ubar, p["P"], Ks, p["Y"], rho)
    # golden printed at 6 significant digits (inference.hpp:176)
    assert np.abs(mean - p["ref"]).max() < 1e-6
    # projected posterior of the information form == projection of the Woodbury form
    R = p["P"] @ Rfull
    Cu = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], R) \
        + (p["sf"] * p["reg"]) ** 2 * p["P"] @ p["P"].T
    mp, cp = orc.post_proj(Cu, p["P"] @ ubar, p["Y"], Ks, rho)
    assert np.abs(mp - p["P"] @ mean).max() < 1e-9 * np.abs(mp).max()
    assert np.abs(cp - p["P"] @ cov @ p["P"].T).max() < 1e-8 * np.abs(cp).max()


def test_woodbury_field_covariance_equals_information_form(orc):
    """The Woodbury posterior covariance D C_u (LikelihoodKOH.ipp:484-488) on the
    whole field equals the information form (C_u^-1 + rho^2 n_o P^T K_s^-1 P)^-1,
    formed here with dense inverses (C_u carries the (sigma_f reg)^2 ridge)."""
    p = laplace_rhs_infer_problem(orc)
    m = p["mesh"]
    lu = orc.LU(p["rp"], p["ci"], p["va"])
    Rfull = lu.solve(np.eye(m.nnodes), trans=True).T
    CuFull = orc.covforcing_project(m.coords, p["covvec"], p["sf"], p["lf"], Rfull) \
        + (p["sf"] * p["reg"]) ** 2 * np.eye(m.nnodes)
    rho, sd, ld = p["hp"]
    Ks = orc.mismatch_cov(p["xs"], sd, ld, p["eps"])
    _, cov = orc.post_woodbury(CuFull, Rfull @ p["F"], p["P"], Ks, p["Y"], rho)
    no = p["Y"].shape[0]
    info = np.linalg.inv(np.linalg.inv(CuFull) + rho * rho * no * p["P"].T @ np.linalg.solve(Ks, p["P"]))
    assert np.abs(cov - info).max() < 1e-8 * np.abs(info).max()
    assert np.abs(cov - cov.T).max() < 1e-12 * np.abs(cov).max()


def test_gp_predict_covariance(orc):
    """predictCovariance (GaussianProcess.hpp:160-190): its diagonal is the
    golden predictive variance, and on well-separated test points it equals
    the inverse of the test block of the joint precision matrix of
    (y(X), f(X*)), the conditional covariance computed another way."""
    hp, X, y, Xp, mean, var = _gp_fixture()
    sg, lg, eg = (float(hp[f"gaussianProcess.{k}"]) for k in ("sigma", "length", "eps"))
    cov = orc.gp_cov(X, sg, lg, eg, Xp, len(y))
    assert np.abs(np.diag(cov) - var).max() < 1e-8
    assert np.abs(cov - cov.T).max() < 1e-15
    Xo = np.array([[0.0], [0.3], [0.55], [0.8], [1.0]])
    Xs = np.array([[0.12], [0.41], [0.68], [0.93]])
    s, l, e = 1.3, 0.25, 0.1
    k = lambda A, B: s * s * np.exp(-(A[:, None, 0] - B[None, :, 0]) ** 2 / (2 * l * l))
    J = np.block([[k(Xo, Xo) + e * e * np.eye(5), k(Xo, Xs)], [k(Xs, Xo), k(Xs, Xs)]])
    ref = np.linalg.inv(np.linalg.inv(J)[5:, 5:])
    assert np.abs(orc.gp_cov(Xo, s, l, e, Xs, 5) - ref).max() < 1e-11


def test_loglik_equals_dense_gaussian_logpdf(orc):
    rng = np.random.default_rng(3)
    ns, no = 7, 4
    A = rng.normal(size=(ns, ns)); Cu = A @ A.T / ns
    xs = rng.uniform(size=(ns, 2))
    Ks = orc.mismatch_cov(xs, 0.2, 0.3, 0.05)
    mean = rng.normal(size=ns)
    Y = rng.normal(size=(no, ns))
    rho = 1.3
    l = orc.loglik(Cu, mean, Y, Ks, rho)
    C = Ks + rho ** 2 * Cu
    sign, logdet = np.linalg.slogdet(C)
    ref = sum(-0.5 * ns * np.log(2 * np.pi) - 0.5 * logdet
              - 0.5 * (y - rho * mean) @ np.linalg.solve(C, y - rho * mean) for y in Y)
    assert abs(l - ref) < 1e-11 * abs(ref)


def test_covforcing_reference_form_equals_dense_triple_product(orc):
    rng = np.random.default_rng(4)
    m = M.grid2d(5)
    cv = orc.body_force(m, fconst=1.0)
    R = rng.normal(size=(3, m.nnodes))
    Cu = orc.covforcing_project(m.coords, cv, 0.1, 0.2, R)
    d = m.coords[:, None, :] - m.coords[None, :, :]
    K = 0.01 * np.exp(-(d ** 2).sum(-1) / (2 * 0.04))
    Cf = cv[:, None] * K * cv[None, :]
    assert np.abs(Cu - R @ Cf @ R.T).max() < 1e-14
    W = rng.normal(size=(m.nnodes, 2))
    assert np.abs(orc.covforcing_apply(m.coords, cv, 0.1, 0.2, W) - Cf @ W).max() < 1e-15
    rows = np.array([0, 7, 12, 35], np.int32)
    assert np.abs(orc.covforcing_rows(m.coords, cv, 0.1, 0.2, W, rows) - (Cf @ W)[rows]).max() < 1e-15


# --------------------------------------------------------------- MCMC
def test_mcmc_gaussian_statistics(orc):
    g = os.path.join(GOLDEN, "mcmcGaussian")
    inp = dict(l.split()[:2] for l in open(os.path.join(g, "input.dat"))
               if l.strip() and not l.startswith("#"))
    s, acc = orc.mcmc_gauss(float(inp["mean"]), float(inp["stdDev"]), float(inp["initialProp"]),
                            float(inp["stdDevProp"]), int(inp["numSamples"]), seed=int(inp["seed"]))
    ref = dict(l.split() for l in open(os.path.join(g, "chain_summary.txt")) if not l.startswith("#"))
    # the recorded chain used Boost's mt19937; only its statistics can be matched
    assert abs(acc - float(ref["acceptance_reported"])) < 0.01
    assert abs(s[:, 0].mean() - 1.0) < 0.03 and abs(s[:, 0].std() - 0.5) < 0.03
    lp = np.log(1 / (np.sqrt(2 * np.pi) * 0.5) * np.exp(-(s[:, 0] - 1) ** 2 / 0.5))
    assert np.abs(s[:, 1] - lp).max() < 1e-12


def test_mcmc_koh_detailed_balance_on_known_target(orc):
    """With Cu = 0, rho enters nowhere except the Jacobian term; the chain must
    then recover the analytic posterior of a 1-sensor GP noise model."""
    ns, no = 1, 200
    rng = np.random.default_rng(5)
    sig_true = 0.3
    Y = rng.normal(scale=sig_true, size=(no, ns))
    xs = np.zeros((1, 1))
    Cu = np.zeros((1, 1))
    mean = np.zeros(1)
    s, acc = orc.mcmc_koh(Cu, mean, Y, xs, 0.0, [0.0, np.log(0.25), 0.0], [0.0, 0.1, 0.0], 20000,
                          prior_mode=2, prior_mu=[0, np.log(0.3), 0], prior_var=[1, 100.0, 1])
    th = s[4000:, 1]
    # Exact posterior of log sigma under the (nearly flat) prior, by quadrature
    grid = np.linspace(np.log(0.2), np.log(0.45), 4001)
    ss = (Y ** 2).sum()
    logp = -no * grid - 0.5 * ss / np.exp(2 * grid) - 0.5 * (grid - np.log(0.3)) ** 2 / 100.0
    w = np.exp(logp - logp.max()); w /= w.sum()
    assert abs(th.mean() - (w * grid).sum()) < 0.01
    assert 0.2 < acc < 0.9


# ------------------------------------------------- SPDE precision, readings
def test_matern_tau_closed_forms(orc):
    # d = 1, nu = 1/2 (exponential kernel): tau^2 = 1 / (2 kappa)
    for kap in (0.5, 3.0, 20.0):
        assert abs(orc.matern_tau(1.0, kap, 0.5, 1) ** 2 - 1.0 / (2 * kap)) < 1e-14 / kap
        # d = 3, nu = 1/2: tau^2 = 1 / (8 pi kappa)
        assert abs(orc.matern_tau(1.0, kap, 0.5, 3) ** 2 - 1.0 / (8 * np.pi * kap)) < 1e-15 / kap
    # sigma scaling
    assert abs(orc.matern_tau(2.0, 3.0, 1.0, 2) - orc.matern_tau(1.0, 3.0, 1.0, 2) / 2.0) < 1e-15


def _matern_1d(orc, n=400, ell=0.05, sigma=0.7, nu=1.5):
    m = M.line1d(n)
    kap = np.sqrt(2 * nu) / ell
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    return m, kap, rp, ci, va, mass


def test_precision_marginal_variance_is_sigma2(orc):
    """The integer-beta SPDE precision Q = tau^2 kappa^4 A M^-1 A of a Matern
    field (nu = 3/2 in 1D) has marginal variance diag(Q^-1) ~ sigma^2 away from
    the boundary; a wrong tau, kappa power or mass placement breaks this."""
    sigma, nu = 0.7, 1.5
    m, kap, rp, ci, va, mass = _matern_1d(orc, sigma=sigma, nu=nu)
    n = m.nnodes
    Q = orc.precision_apply(rp, ci, va, mass, sigma, kap, nu, 1, np.eye(n))
    assert np.abs(Q - Q.T).max() < 1e-9 * np.abs(Q).max()
    var = np.diag(np.linalg.inv(Q))
    mid = var[n // 2 - 40:n // 2 + 40]
    assert np.all(np.abs(mid - sigma ** 2) < 0.03 * sigma ** 2), (mid.min(), mid.max())


def test_typical_readings_statistics(orc):
    rng = np.random.default_rng(9)
    ns = 5
    xs = rng.uniform(size=(ns, 2))
    A = rng.normal(size=(ns, ns)) * 0.1
    Cu = A @ A.T
    mean = rng.normal(size=ns)
    rho, sd, ld, eps = 1.3, 0.2, 0.3, 0.05
    Y = orc.typical_readings(Cu, mean, xs, eps, rho, sd, ld, 40000, seed=4)
    C = orc.mismatch_cov(xs, sd, ld, eps) + rho ** 2 * Cu
    assert np.abs(Y.mean(0) - rho * mean).max() < 4 * np.sqrt(np.diag(C).max() / 40000)
    assert np.abs(np.cov(Y.T) - C).max() < 0.03 * np.abs(C).max()


def test_mlp_oracle_gradient_by_finite_differences():
    from oracle import mlp
    rng = np.random.default_rng(12)
    p = rng.normal(size=mlp.NP) * 0.2
    X = rng.normal(size=(7, 3))
    t = rng.normal(size=7)
    _, g = mlp.loss_grad(p, X, t)
    for idx in list(range(0, 5)) + [200, 300, 4000, 4400, 4480]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (mlp.loss_grad(pp, X, t)[0] - mlp.loss_grad(pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[idx]) < 1e-7 + 1e-6 * abs(g[idx])


def test_ess_of_ar1_chains_matches_closed_form():
    """AR(1) with coefficient a has integrated autocorrelation time
    (1 + a) / (1 - a); independent draws have ESS = number of draws."""
    from statfem_amd import diagnostics as D
    rng = np.random.default_rng(21)
    m, n, a = 8, 20000, 0.8
    x = np.zeros((m, n))
    e = rng.normal(size=(m, n))
    for t in range(1, n):
        x[:, t] = a * x[:, t - 1] + e[:, t]
    expect = m * n * (1 - a) / (1 + a)
    assert abs(D.ess(x) / expect - 1) < 0.1
    iid = rng.normal(size=(m, n))
    assert abs(D.ess(iid) / (m * n) - 1) < 0.1
    assert D.split_rhat(iid) < 1.01
    shifted = iid + np.arange(m)[:, None]
    assert D.split_rhat(shifted) > 2


# ----------------------------------------------- conductivity (LHS) uncertainty
def test_laplace_inverse_targets_golden(orc):
    """The oracle's laplaceInverse log target against the recorded chain
    (every 10th row of logMLsamples.dat: natural hyperparameters and their log
    target, 6 significant digits). The reference evaluates the likelihood at
    the FE solution of the map's own conductivities (inference.hpp:139; the
    chain never re-solves): that reproduces every row to print precision,
    while re-solving at each row's conductivities does not."""
    from . import lhs_deck
    d = lhs_deck.load()
    m = d["mesh"]
    P = lhs_deck.projection(d, m.nnodes)
    u, _ = orc.lhs_solution(m, d["bc"], d["anchors"], d["w0"], *d["gp"], nq1d=3)
    mean = P @ u
    dev = [lhs_deck.oracle_target(orc, d, np.log(r[:6]), mean) - r[6] for r in d["golden"]]
    assert len(dev) == 350
    assert np.abs(dev).max() < 1e-3
    r = d["golden"][100]
    ur, _ = orc.lhs_solution(m, d["bc"], d["anchors"], r[:5], *d["gp"], nq1d=3)
    assert abs(lhs_deck.oracle_target(orc, d, np.log(r[:6]), P @ ur) - r[6]) > 10.0


def test_laplace_inverse_conductivity_interpolates_anchors(orc):
    """The GP-mean conductivity passes (to the nugget) through the anchor
    values and is linear in them."""
    m = M.line1d(400)
    A = np.array([[0.0], [0.25], [0.5], [0.75], [1.0]])
    w = np.array([0.7, 1.0, 0.7, 0.4, 0.7])
    k = orc.lhs_conductivity(m, A, w, 1.0, 0.32, 5e-5)
    cen = m.coords[m.conn].mean(axis=1)[:, 0]
    for a, v in zip(A[:, 0], w):
        e = np.argmin(np.abs(cen - a))
        assert abs(k[e] - v) < 0.02
    k2 = orc.lhs_conductivity(m, A, 2.0 * w + 1.0, 1.0, 0.32, 5e-5)
    k1 = orc.lhs_conductivity(m, A, np.ones(5), 1.0, 0.32, 5e-5)
    assert np.allclose(k2, 2.0 * k + k1, rtol=0, atol=1e-12)


def test_lhs_draws_counter_layout(orc):
    """Each pair of proposal increments is the Box-Muller pair of Philox block
    {step, chain, 32 + i, 0}; the uniform is block {step, chain, 31, 0}."""
    seed, chain, step = 0x1234567890ABCDEF, 3, 17
    z, u = orc.lhs_draws(seed, chain, step, 6)
    key = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)

    def unit(hi, lo):
        return ((int(hi) << 32 | int(lo)) >> 11) * 2.0 ** -53 + 2.0 ** -54

    for i in range(3):
        a = orc.philox(np.array([step, chain, 32 + i, 0], np.uint32), key)
        r = np.sqrt(-2 * np.log(unit(a[0], a[1])))
        t = 2 * np.pi * unit(a[2], a[3])
        assert abs(z[2 * i] - r * np.cos(t)) < 1e-14 and abs(z[2 * i + 1] - r * np.sin(t)) < 1e-14
    a = orc.philox(np.array([step, chain, 31, 0], np.uint32), key)
    assert u == unit(a[0], a[1])
    zs = np.array([orc.lhs_draws(7, c, t, 6)[0] for c in range(20) for t in range(100)])
    assert abs(zs.mean()) < 0.05 and abs(zs.std() - 1.0) < 0.05


# ------------------------------------------ sparse-precision GP (GaussianProcessQ)
def _gpq_case(orc, dim=2):
    m = M.grid2d(9, 7) if dim == 2 else M.line1d(30)
    nu = 1.0 if dim == 2 else 1.5
    kap = np.sqrt(2 * nu) / 0.3
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    Q = orc.gpq_precision_dense(rp, ci, va, mass, 0.7, kap, nu, dim)
    rng = np.random.default_rng(dim)
    ns, no = 11, 4
    P = np.zeros((ns, m.nnodes))
    for i, j in enumerate(rng.choice(m.nnodes, ns, replace=False)):
        P[i, j] = 1.0
    P[3, :] = 0.0                     # one sensor interpolating between two nodes
    P[3, [1, 2]] = [0.4, 0.6]
    Y = 0.3 * rng.standard_normal((ns, no))
    return Q, P, Y


@pytest.mark.parametrize("dim", [1, 2])
def test_gpq_loglik_is_the_joint_gaussian_density(orc, dim):
    """The oracle's GaussianProcessQ log-likelihood (two Cholesky log-dets as
    the reference) equals the density of the stacked readings under the model
    y_i = P u + e_i with one shared u ~ N(0, Q^-1): covariance J (x) P Q^-1 P^T +
    sigma_e^2 I, evaluated by brute force."""
    Q, P, Y = _gpq_case(orc, dim)
    ns, no = Y.shape
    G = P @ np.linalg.inv(Q) @ P.T
    for eps in (0.05, 0.2):
        K = np.kron(np.ones((no, no)), G) + eps * eps * np.eye(ns * no)
        yv = Y.T.reshape(-1)
        sign, ld = np.linalg.slogdet(K)
        ref = -0.5 * (yv @ np.linalg.solve(K, yv) + ld + ns * no * np.log(2 * np.pi))
        assert sign > 0
        assert abs(orc.gpq_loglik(Q, P, Y, eps) - ref) <= 1e-10 * abs(ref)


def test_gpq_posterior_is_the_gaussian_conditional(orc):
    """predictMean / predictCovariance against the conditional of the joint
    Gaussian (u, y_1 .. y_no) by brute force."""
    Q, P, Y = _gpq_case(orc, 2)
    ns, no = Y.shape
    eps = 0.1
    Ci = np.linalg.inv(Q)
    Pb = np.vstack([P] * no)
    K = Pb @ Ci @ Pb.T + eps * eps * np.eye(ns * no)
    mean_ref = Ci @ Pb.T @ np.linalg.solve(K, Y.T.reshape(-1))
    cov_ref = Ci - Ci @ Pb.T @ np.linalg.solve(K, Pb @ Ci)
    mean, cov = orc.gpq_predict(Q, P, Y, eps)
    assert np.abs(mean - mean_ref).max() <= 1e-10 * np.abs(mean_ref).max()
    assert np.abs(cov - cov_ref).max() <= 1e-10 * np.abs(cov_ref).max()


def _gpq_frac_case(orc, nu, dim=2, degree=4):
    from statfem_amd import rational
    from scipy.linalg import eigh
    m = M.grid2d(8, 7) if dim == 2 else M.line1d(40)
    kap = np.sqrt(2 * nu) / 0.3
    rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    beta = (nu + dim / 2) / 2
    alpha = max(1.0, np.floor(beta))
    top, bot, _ = rational.brasil_factors(beta - alpha, degree)
    n = m.nnodes
    A = orc.spmv(rp, ci, va, np.eye(n))
    lam, V = eigh(A, np.diag(mass))            # A V = M V diag(lam), V^T M V = I
    return m, kap, rp, ci, va, mass, beta, alpha, top, bot, lam, V


@pytest.mark.parametrize("nu,dim", [(0.5, 2), (1.5, 2), (1.0, 1)])
def test_gpq_fractional_factors_are_spectral_polynomials(orc, nu, dim):
    """P_l and P_r of the oracle against the generalized eigenvectors of
    (A, M): lead prod_j (I - r_j M^-1 A) = V diag(lead prod_j (1 - r_j lam)) V^T M."""
    m, kap, rp, ci, va, mass, beta, alpha, top, bot, lam, V = _gpq_frac_case(orc, nu, dim)
    for roots in (top, bot):
        ref = V @ np.diag(roots[0] * np.prod(1.0 - np.outer(lam, roots[1:]), axis=1)) @ V.T @ np.diag(mass)
        got = orc.gpq_frac_factor(rp, ci, va, mass, roots)
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("nu,dim", [(0.5, 2), (1.5, 2), (1.0, 1)])
def test_gpq_fractional_covariance_is_the_spectral_power(orc, nu, dim):
    """The field covariance P_r Q_v^-1 P_r^T of the oracle's fractional precision
    is (1/c) V diag(r(1/lam)^2 lam^(-2 alpha)) V^T exactly (r the factor form
    a prod (x - c_i) / (b prod (x - d_j))) and, within the rational
    approximation's error, the fractional Matern covariance (1/c) V lam^(-2 beta) V^T."""
    from statfem_amd import rational
    m, kap, rp, ci, va, mass, beta, alpha, top, bot, lam, V = _gpq_frac_case(orc, nu, dim)
    sigma = 0.7
    Qv, Pr = orc.gpq_precision_frac_dense(rp, ci, va, mass, sigma, kap, nu, dim, top, bot)
    C = Pr @ np.linalg.solve(Qv, Pr.T)
    c = orc.matern_tau(sigma, kap, nu, dim) ** 2 * kap ** (4 * beta)
    rr = rational.factor_form(top, bot, 1.0 / lam)
    exact = V @ np.diag(rr ** 2 * lam ** (-2 * alpha)) @ V.T / c
    # (Q_v's condition is cond(Q) times cond(P_l)^2, ~1e10 in 1D here)
    assert np.abs(C - exact).max() <= 1e-6 * np.abs(exact).max()
    frac = V @ np.diag(lam ** (-2 * beta)) @ V.T / c
    x = 1.0 / lam
    relerr = np.abs(rr / x ** (beta - alpha) - 1.0).max()
    assert relerr < 0.05
    assert np.abs(C - frac).max() <= 2.5 * relerr * np.abs(frac).max()
    # the integer part alone is off by the missing power
    assert np.abs(V @ np.diag(lam ** (-2 * alpha)) @ V.T / c - frac).max() > 10 * np.abs(C - frac).max()


def test_gpq_fractional_loglik_and_posterior_are_the_joint_gaussian(orc):
    """With u = P_r v, v ~ N(0, Q_v^-1): the reference's two-log-determinant
    likelihood and its P_r-projected posterior against the joint density and
    the Gaussian conditional of (u, y_1 .. y_no) by brute force."""
    m, kap, rp, ci, va, mass, beta, alpha, top, bot, lam, V = _gpq_frac_case(orc, 0.5)
    Qv, Pr = orc.gpq_precision_frac_dense(rp, ci, va, mass, 0.7, kap, 0.5, 2, top, bot)
    rng = np.random.default_rng(5)
    ns, no = 9, 3
    P = np.zeros((ns, m.nnodes))
    for i, j in enumerate(rng.choice(m.nnodes, ns, replace=False)):
        P[i, j] = 1.0
    Y = 0.3 * rng.standard_normal((ns, no))
    Cu = Pr @ np.linalg.solve(Qv, Pr.T)
    Pb = np.vstack([P] * no)
    yv = Y.T.reshape(-1)
    for eps in (0.05, 0.2):
        K = Pb @ Cu @ Pb.T + eps * eps * np.eye(ns * no)
        _, ld = np.linalg.slogdet(K)
        ref = -0.5 * (yv @ np.linalg.solve(K, yv) + ld + ns * no * np.log(2 * np.pi))
        assert abs(orc.gpq_loglik(Qv, P, Y, eps, Pr) - ref) <= 1e-9 * abs(ref)
    eps = 0.1
    K = Pb @ Cu @ Pb.T + eps * eps * np.eye(ns * no)
    mean_ref = Cu @ Pb.T @ np.linalg.solve(K, yv)
    cov_ref = Cu - Cu @ Pb.T @ np.linalg.solve(K, Pb @ Cu)
    mean, cov = orc.gpq_predict(Qv, P, Y, eps, Pr)
    assert np.abs(mean - mean_ref).max() <= 1e-9 * np.abs(mean_ref).max()
    assert np.abs(cov - cov_ref).max() <= 1e-9 * np.abs(cov_ref).max()


def test_precision_beta2_is_matern_7_2(orc):
    """beta = 2 (nu = 7/2 in 1D, Q = c (A M^-1)^3 A): the covariance Q^-1 is the
    Matern-7/2 kernel sigma^2 e^-a (1 + a + 2a^2/5 + a^3/15), a = sqrt(7) r / ell,
    at the interior (variance and three lags)."""
    sigma, nu, ell = 0.7, 3.5, 0.05
    m, kap, rp, ci, va, mass = _matern_1d(orc, n=400, ell=ell, sigma=sigma, nu=nu)
    Q = orc.precision_apply(rp, ci, va, mass, sigma, kap, nu, 1, np.eye(m.nnodes))
    Cm = np.linalg.inv(Q)
    c = m.nnodes // 2
    x = m.coords[:, 0]
    for k in (0, 10, 20, 40):
        a = np.sqrt(7) * abs(x[c + k] - x[c]) / ell
        ref = sigma ** 2 * np.exp(-a) * (1 + a + 0.4 * a * a + a ** 3 / 15)
        assert abs(Cm[c, c + k] - ref) < 0.003 * sigma ** 2, (k, Cm[c, c + k], ref)


# --------------------------------------------- first-order perturbation (ForwardPerturbation)
def _fpert_case(orc, nx=5, ny=4):
    m = M.grid2d(nx, ny)
    bc = M.dirichlet_all_boundary(m)
    kap = 1.0 + 0.3 * np.sin(np.arange(m.nel))
    F = orc.body_force(m, nq1d=2, fconst=1.0)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, kap, nq1d=2, bc_dofs=bc, F=F)
    # stiffness derivative of each element's conductivity: the unit-conductivity
    # element matrix on its nodes, zero on constrained rows (row-zero form)
    Ke = orc.element_matrices(m, orc.POISSON, np.ones(m.nel), nq1d=2)
    mask = np.zeros(m.nnodes, bool)
    mask[bc] = True
    ddofs = m.conn.astype(np.int32)
    dblock = np.array([np.where(mask[m.conn[e]][:, None], 0.0, Ke[e]) for e in range(m.nel)])
    cen = m.coords[m.conn].mean(axis=1)
    rng = np.random.default_rng(2)
    ns = 6
    P = np.zeros((ns, m.nnodes))
    P[np.arange(ns), rng.choice(np.setdiff1d(np.arange(m.nnodes), bc), ns, replace=False)] = 1.0
    cv = orc.body_force(m, nq1d=2, fconst=1.0)
    cv[bc] = 0.0
    return m, bc, kap, rp, ci, va, F, P, ddofs, dblock, cen, cv


def test_fpert_material_term_is_the_linearised_covariance(orc):
    """Without forcing covariance, the material part is the first-order
    covariance of P u(kappa) under kappa ~ N(kappa_bar, C_k): J C_k J^T with
    J = dPu/dkappa from central finite differences of the forward solve."""
    m, bc, kap, rp, ci, va, F, P, ddofs, dblock, cen, cv = _fpert_case(orc)
    mean, cov, u = orc.fpert_update(rp, ci, va, F, m.coords, cv, 0.0, 0.3, P, ddofs, dblock, cen, 0.2, 0.4,
                                    reg_factor=0.0)
    J = np.zeros((P.shape[0], m.nel))
    f0 = orc.body_force(m, nq1d=2, fconst=1.0)
    for e in range(m.nel):
        us = []
        for sgn in (1, -1):
            k2 = kap.copy()
            k2[e] += sgn * 1e-4
            r2, c2, v2, F2 = orc.assemble(m, orc.POISSON, k2, nq1d=2, bc_dofs=bc, F=f0.copy())
            us.append(P @ orc.LU(r2, c2, v2).solve(F2))
        J[:, e] = (us[0] - us[1]) / 2e-4
    Ck = orc.mismatch_cov(cen, 0.2, 0.4, 0.0)
    ref = J @ Ck @ J.T
    assert np.abs(cov - ref).max() <= 1e-6 * np.abs(ref).max()
    assert np.abs(mean - P @ u).max() == 0.0


def test_fpert_forcing_terms_by_factorised_forcing(orc):
    """The forcing parts through a factor of C_f (C_f = L L^T, A^-1 L by direct
    solves): R C_f R^T + sum C_k R D_e A^-1 L (R D_d A^-1 L)^T. Distinguishes A^-1
    from A^-T (the row-zero operator is not symmetric)."""
    m, bc, kap, rp, ci, va, F, P, ddofs, dblock, cen, cv = _fpert_case(orc)
    sf, lf, sk, lk = 0.5, 0.3, 0.2, 0.4
    _, cov, u = orc.fpert_update(rp, ci, va, F, m.coords, cv, sf, lf, P, ddofs, dblock, cen, sk, lk, reg_factor=0.0)
    N = m.nnodes
    Cf = orc.covforcing_apply(m.coords, cv, sf, lf, np.eye(N))
    w, V = np.linalg.eigh(0.5 * (Cf + Cf.T))
    L = V * np.sqrt(np.clip(w, 0, None))
    lu = orc.LU(rp, ci, va)
    AL = lu.solve(L)
    R = lu.solve(np.ascontiguousarray(P.T), trans=True).T
    Ck = orc.mismatch_cov(cen, sk, lk, 0.0)
    B = []
    for e in range(m.nel):
        De = np.zeros((N, N))
        De[np.ix_(ddofs[e], ddofs[e])] += dblock[e]
        B.append(np.hstack([R @ De @ AL, (R @ De @ u)[:, None]]))
    ref = (R @ L) @ (R @ L).T
    for e in range(m.nel):
        for d in range(m.nel):
            ref += Ck[e, d] * B[e] @ B[d].T
    assert np.abs(cov - ref).max() <= 1e-10 * np.abs(ref).max()


def test_threaded_banded_solve_matches_serial(orc):
    """The CPU baseline's std::thread variant of the banded Cholesky solve
    (columns split over threads) returns the serial result exactly."""
    m = M.grid2d(30)
    bc = M.dirichlet_all_boundary(m)
    F = orc.body_force(m, nq1d=2, fconst=1.0)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, 1.0 + 0.1 * np.arange(m.nel) % 3, nq1d=2, bc_dofs=bc, F=F)
    mask = np.zeros(m.nnodes, np.uint8)
    mask[bc] = 1
    ch = orc.Chol(rp, ci, va, mask)
    B = np.random.default_rng(3).standard_normal((m.nnodes, 11))
    B[bc] = 0.0
    assert np.array_equal(ch.solve(B), ch.solve(B, threads=4))


def test_forward_gauss_2d_composition_matches_reference_form(orc):
    """oracle.forward (banded Cholesky, restored constrained components,
    separable C_f) against the reference form at a small size: transposed LU
    solves and the dense C_f projection (CovarianceForcing.ipp:37-109)."""
    from oracle.forward import ForwardGauss2D
    from statfem_amd import configs
    cfg = configs.c2_inputs(seed=4, n=14, ns=9)
    m = cfg["mesh"]
    fg = ForwardGauss2D(m, cfg["bc"], cfg["kappa"], threads=2)
    ref = fg.forward_gauss(cfg["cols"], cfg["vals"], 0.1, 0.2, 1e-3, sample=[0, 3, 8])
    P = M.dense_projection(m.nnodes, cfg["cols"], cfg["vals"])
    lu = orc.LU(fg.rp, fg.ci, fg.va)
    R = lu.solve(P.T.copy(), trans=True)
    assert np.abs(ref["Rt"] - R[:, [0, 3, 8]]).max() <= 1e-12 * np.abs(R).max()
    Cu = orc.covforcing_project(m.coords, fg.cv, 0.1, 0.2, R.T.copy()) + (1e-4) ** 2 * P @ P.T
    assert np.abs(ref["Cu"] - Cu[np.ix_([0, 3, 8], [0, 3, 8])]).max() <= 1e-12 * np.abs(Cu).max()
    mean = R.T @ fg.F
    assert np.abs(ref["mean"] - mean).max() <= 1e-12 * np.abs(mean).max()


def test_elasticity_spde_reference_matches_dense_composition(orc):
    """oracle.forward.elasticity_spde_reference (banded Cholesky factors, sparse
    products) against the dense composition of the same oracle operators at 4^3:
    C_u = R B Q^-1 B^T R^T, the field block C_u P^T and the Woodbury field mean
    of or_post_woodbury."""
    from oracle.forward import elasticity_spde_reference
    from statfem_amd import configs
    cfg = configs.c3_inputs(seed=3, n=4, ns=6)
    m = cfg["mesh"]
    y = np.random.default_rng(0).normal(size=6) * 1e-3
    ref = elasticity_spde_reference(cfg, y)
    F = orc.body_force(m, nq1d=2, dofpn=3, fconst=[0.0, 0.0, cfg["fbar"]])
    rp, ci, va, F = orc.assemble(m, orc.ELASTICITY, cfg["E"], cfg["nu"], nq1d=2, bc_dofs=cfg["bc"], F=F)
    Ae = orc.LU(rp, ci, va)
    kap = cfg["kappa"]
    rpm, cim, vam, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
    Am = _dense(rpm, cim, vam, m.nnodes)
    mass = orc.body_force(m, nq1d=2, fconst=1.0)
    tau = orc.matern_tau(cfg["sigma_g"], kap, cfg["nu_m"], 3)
    Ainv = np.linalg.inv(Am)
    Cg = Ainv @ np.diag(mass) @ Ainv / (tau * kap ** 2) ** 2
    w = mass.copy()
    w[cfg["bottom"]] = 0.0
    N = 3 * m.nnodes
    B = np.zeros((N, m.nnodes))
    B[3 * np.arange(m.nnodes) + 2, np.arange(m.nnodes)] = w
    P = np.zeros((len(cfg["cols"]), N))
    for s in range(P.shape[0]):
        np.add.at(P[s], cfg["cols"][s], cfg["vals"][s])
    Rt = Ae.solve(P.T.copy(), trans=True)
    Cu = Rt.T @ B @ Cg @ B.T @ Rt
    ubar = Ae.solve(F)
    CuFull = Ae.solve(B @ Cg @ B.T @ Ae.solve(np.eye(N), trans=True), trans=False)
    Ks = orc.mismatch_cov(cfg["xs"], 0.0, 1.0, cfg["sigma_e"])
    fm, _ = orc.post_woodbury(CuFull, ubar, P, Ks, np.atleast_2d(y), 1.0, want_cov=False)
    assert np.abs(ref["Cu"] - Cu).max() <= 1e-12 * np.abs(Cu).max()
    assert np.abs(ref["mean"] - P @ ubar).max() <= 1e-12 * np.abs(P @ ubar).max()
    assert np.abs(ref["field"] - CuFull @ P.T).max() <= 1e-12 * np.abs(CuFull @ P.T).max()
    assert np.abs(ref["field_mean"] - fm).max() <= 1e-12 * np.abs(fm).max()
