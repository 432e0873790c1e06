"""Batched uneven problems: one block-diagonal system, per-problem PCG scalars,
against the oracle problem by problem."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_forward_gauss_matches_oracle_per_problem(ctx, orc):
    from statfem_amd import configs, hip, mesh as M
    from statfem_amd.batch import BatchedForwardGauss
    probs = configs.c4_inputs(seed=5, nprob=5, nmin=4, nmax=14, smin=2, smax=9)
    bf = BatchedForwardGauss(ctx, probs)
    means, Cus, st = bf.updateSolution(hip.default_opts(tol=1e-12))
    for p, mean, Cu in zip(probs, means, Cus):
        m = p["mesh"]
        F = orc.body_force(m, nq1d=2, fconst=1.0)
        rp, ci, va, F = orc.assemble(m, orc.POISSON, 1.0, nq1d=2, bc_dofs=p["bc"], F=F)
        P = M.dense_projection(m.nnodes, p["cols"], p["vals"])
        R = orc.LU(rp, ci, va).solve(P.T.copy(), trans=True).T.copy()
        cv = orc.body_force(m, nq1d=2, fconst=1.0)
        cv[p["bc"]] = 0.0
        Cu_ref = orc.covforcing_project(m.coords, cv, p["sigma_f"], p["ell_f"], R) \
            + (p["sigma_f"] * p["reg"]) ** 2 * P @ P.T
        mean_ref = R @ F
        assert np.abs(mean - mean_ref).max() <= 1e-9 * np.abs(mean_ref).max()
        assert np.abs(Cu - Cu_ref).max() <= 1e-8 * np.abs(Cu_ref).max()


def test_batched_solve_forward_and_transposed(ctx, orc):
    """Per-problem relative residual targets: problems whose right-hand sides
    differ in scale by 1e6 still each reach the tolerance."""
    from statfem_amd import configs, hip, mesh as M
    probs = configs.c4_inputs(seed=6, nprob=4, nmin=3, nmax=11, smin=1, smax=3)
    mu, off = M.union([p["mesh"] for p in probs])
    s = hip.SystemSolveHip(ctx, mu)
    kap = np.exp(np.random.default_rng(2).normal(size=mu.nel))
    s.assemble(hip.POISSON, kap, nq1d=2)
    bc = np.concatenate([p["bc"] + off[q] for q, p in enumerate(probs)]).astype(np.int32)
    s.applyConstraints(bc)
    rp, ci, va, _ = orc.assemble(mu, orc.POISSON, kap, nq1d=2, bc_dofs=bc)
    lu = orc.LU(rp, ci, va)
    rng = np.random.default_rng(3)
    k = 6
    B = rng.normal(size=(mu.nnodes, k))
    for q in range(len(probs)):
        B[off[q]:off[q + 1]] *= 10.0 ** (2 * q)
    B[off[1]:off[2], 3:] = 0.0      # a problem with fewer right-hand sides
    seg = np.ascontiguousarray(off, dtype=np.int32)
    for trans in (0, 1):
        X = ctx.empty((mu.nnodes, k))
        st = hip.SolveStats()
        o = hip.default_opts(tol=1e-11)
        Bd = ctx.to_device(B)
        hip._check(hip.lib().sfh_solve_batched(s.h, len(probs), hip._p(seg), trans, k, Bd.ptr, X.ptr, C.byref(o),
                                               C.byref(st)))
        Xh = X.download()
        ref = lu.solve(B, trans=bool(trans))
        for q in range(len(probs)):
            a, b = off[q], off[q + 1]
            scale = np.abs(ref[a:b]).max()
            assert np.abs(Xh[a:b] - ref[a:b]).max() <= 1e-9 * scale
        assert np.all(Xh[off[1]:off[2], 3:] == 0.0)


def test_batched_multigrid_on_declared_grids(ctx, orc):
    """A batch of declared grids of different sizes (odd and even node counts,
    one small enough for a single coarse level) runs the multigrid-
    preconditioned solve in shared launches; every problem matches its own
    direct solve, whatever the scale of its right-hand sides."""
    from statfem_amd import hip, mesh as M
    sizes = [(40, 40), (9, 9), (57, 31), (24, 24)]
    meshes = [M.grid2d(a, b) for a, b in sizes]
    mu, off = M.union(meshes)
    s = hip.SystemSolveHip(ctx, mu)
    kap = np.exp(0.6 * np.random.default_rng(4).normal(size=mu.nel))
    s.assemble(hip.POISSON, kap, nq1d=2)
    bc = np.concatenate([M.dirichlet_all_boundary(m) + off[q] for q, m in enumerate(meshes)]).astype(np.int32)
    s.applyConstraints(bc)
    seg = np.ascontiguousarray(off, dtype=np.int32)
    nx = np.array([a + 1 for a, b in sizes], np.int32)
    ny = np.array([b + 1 for a, b in sizes], np.int32)
    hip._check(hip.lib().sfh_system_set_segment_grids(s.h, len(sizes), hip._p(seg), hip._p(nx), hip._p(ny)))
    rp, ci, va, _ = orc.assemble(mu, orc.POISSON, kap, nq1d=2, bc_dofs=bc)
    lu = orc.LU(rp, ci, va)
    rng = np.random.default_rng(5)
    k = 20
    B = rng.normal(size=(mu.nnodes, k))
    for q in range(len(sizes)):
        B[off[q]:off[q + 1]] *= 10.0 ** (2 * q)
    B[off[2]:off[3], 7] = 0.0
    for trans in (0, 1):
        X = ctx.empty((mu.nnodes, k))
        st = hip.SolveStats()
        o = hip.default_opts(tol=1e-11)
        Bd = ctx.to_device(B)
        hip._check(hip.lib().sfh_solve_batched(s.h, len(sizes), hip._p(seg), trans, k, Bd.ptr, X.ptr, C.byref(o),
                                               C.byref(st)))
        assert st.precond == 2 and st.layout == 2 and st.converged == 1, st.as_dict()
        assert st.inner_iters < 40, st.as_dict()
        Xh = X.download()
        ref = lu.solve(B, trans=bool(trans))
        for q in range(len(sizes)):
            a, b = off[q], off[q + 1]
            scale = np.abs(ref[a:b]).max()
            assert np.abs(Xh[a:b] - ref[a:b]).max() <= 1e-9 * scale, (q, trans)
        assert np.all(Xh[off[2]:off[3], 7] == 0.0)


def test_batched_forcing_covariance_equals_per_problem(ctx):
    """sfh_covforcing_batched over a union of grids gives what the per-grid
    sfh_covforcing_grid gives on each block."""
    from statfem_amd import hip, mesh as M
    sizes = [(12, 9), (20, 20), (7, 15)]
    meshes = [M.grid2d(a, b) for a, b in sizes]
    mu, off = M.union(meshes)
    rng = np.random.default_rng(14)
    k = 5
    W = rng.normal(size=(mu.nnodes, k))
    cv = rng.uniform(0.5, 1.5, size=mu.nnodes)
    sig = np.array([0.3, 1.1, 0.7])
    ell = np.array([0.1, 0.25, 0.4])
    nx = np.array([a + 1 for a, b in sizes], np.int32)
    ny = np.array([b + 1 for a, b in sizes], np.int32)
    axes = np.concatenate([np.r_[m.grid[0], m.grid[1]] for m in meshes])
    seg = np.ascontiguousarray(off, dtype=np.int32)
    Wd, cvd, V = ctx.to_device(W), ctx.to_device(cv), ctx.empty((mu.nnodes, k))
    hip._check(hip.lib().sfh_covforcing_batched(ctx.h, 3, hip._p(seg), hip._p(nx), hip._p(ny), hip._p(axes),
                                                cvd.ptr, hip._p(sig), hip._p(ell), k, Wd.ptr, V.ptr))
    Vb = V.download()
    for q, m in enumerate(meshes):
        a, b = off[q], off[q + 1]
        Wq, cq, Vq = ctx.to_device(W[a:b]), ctx.to_device(cv[a:b]), ctx.empty((b - a, k))
        ax, ay = np.ascontiguousarray(m.grid[0]), np.ascontiguousarray(m.grid[1])
        hip._check(hip.lib().sfh_covforcing_grid(ctx.h, len(ax), len(ay), 1, hip._p(ax), hip._p(ay), None, cq.ptr,
                                                 sig[q], ell[q], k, Wq.ptr, Vq.ptr))
        assert np.array_equal(Vb[a:b], Vq.download())


def test_batched_forcing_covariance_with_a_line_problem(ctx):
    """A batch mixing a 2D grid and a single row of nodes (ny = 1, the 1D
    product): each block equals sfh_covforcing_grid on that grid alone, bit
    for bit."""
    from statfem_amd import hip
    rng = np.random.default_rng(21)
    shapes = [(9, 1), (6, 11), (17, 1)]
    axes_l = [(np.sort(rng.uniform(0, 1, a)), np.sort(rng.uniform(0, 1, b)) if b > 1 else np.zeros(1))
              for a, b in shapes]
    off = np.r_[0, np.cumsum([a * b for a, b in shapes])].astype(np.int32)
    k = 6
    N = int(off[-1])
    W = rng.normal(size=(N, k))
    cv = rng.uniform(0.5, 1.5, size=N)
    sig = np.array([0.4, 1.3, 0.9])
    ell = np.array([0.2, 0.15, 0.5])
    nx = np.array([a for a, b in shapes], np.int32)
    ny = np.array([b for a, b in shapes], np.int32)
    axes = np.concatenate([np.r_[ax, ay] for ax, ay in axes_l])
    Wd, cvd, V = ctx.to_device(W), ctx.to_device(cv), ctx.empty((N, k))
    hip._check(hip.lib().sfh_covforcing_batched(ctx.h, 3, hip._p(off), hip._p(nx), hip._p(ny), hip._p(axes),
                                                cvd.ptr, hip._p(sig), hip._p(ell), k, Wd.ptr, V.ptr))
    Vb = V.download()
    for q, (ax, ay) in enumerate(axes_l):
        a, b = off[q], off[q + 1]
        Wq, cq, Vq = ctx.to_device(W[a:b]), ctx.to_device(cv[a:b]), ctx.empty((b - a, k))
        ax, ay = np.ascontiguousarray(ax), np.ascontiguousarray(ay)
        hip._check(hip.lib().sfh_covforcing_grid(ctx.h, len(ax), int(ny[q]), 1, hip._p(ax), hip._p(ay), None,
                                                 cq.ptr, sig[q], ell[q], k, Wq.ptr, Vq.ptr))
        assert np.array_equal(Vb[a:b], Vq.download()), q
        # and the separable kernel itself, from its definition
        X = np.array([(x, y) for y in ay for x in ax])
        d2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
        ref = sig[q] ** 2 * cv[a:b, None] * (np.exp(-d2 / (2 * ell[q] ** 2)) @ (cv[a:b, None] * W[a:b]))
        assert np.abs(Vb[a:b] - ref).max() <= 1e-12 * np.abs(ref).max(), q


@pytest.mark.parametrize("m,n", [(7, 5), (127, 127), (127, 1)])
def test_gram_segmented_matches_per_problem(ctx, m, n):
    """sfh_gram_segmented (all problems in one split-K launch) against X_q^T V_q
    per problem, with an empty problem (its block is zero) and problems from
    a few rows to tens of thousands."""
    from statfem_amd import hip
    rng = np.random.default_rng(m + n)
    sizes = [40, 0, 25000, 3, 9000, 1]
    seg = np.r_[0, np.cumsum(sizes)].astype(np.int32)
    N = int(seg[-1])
    X = rng.normal(size=(N, m))
    V = rng.normal(size=(N, n))
    Xd, Vd, G = ctx.to_device(X), ctx.to_device(V), ctx.empty((len(sizes), m, n))
    hip._check(hip.lib().sfh_gram_segmented(ctx.h, len(sizes), hip._p(seg), m, n, Xd.ptr, m, Vd.ptr, n, G.ptr))
    Gh = G.download()
    for q in range(len(sizes)):
        a, b = seg[q], seg[q + 1]
        ref = X[a:b].T @ V[a:b]
        if b == a:
            assert np.all(Gh[q] == 0.0)
            continue
        scale = np.sqrt((X[a:b] ** 2).sum(0))[:, None] * np.sqrt((V[a:b] ** 2).sum(0))[None, :]
        assert np.all(np.abs(Gh[q] - ref) <= 1e-13 * scale), q


@pytest.mark.parametrize("k,precond", [(1, 0), (6, 0), (6, 1), (3, 1)])
def test_narrow_batch_matches_oracle(ctx, orc, k, precond):
    """A batch of three grids (two with a multigrid hierarchy) solved with
    1 to 6 columns per problem: 8-column tiles (2 lanes per node) for the
    multigrid and the Jacobi stencil kernels (Jacobi with 3 columns: CSR);
    every problem against the oracle's direct solve."""
    from statfem_amd import hip, mesh as M
    ms = [M.grid2d(24, 17), M.grid2d(9, 7), M.grid2d(31, 12)]
    um, off = M.union(ms)
    seg = np.ascontiguousarray(off, np.int32)
    s = hip.SystemSolveHip(ctx, um)
    nx = np.array([len(m.grid[0]) for m in ms], np.int32)
    ny = np.array([len(m.grid[1]) for m in ms], np.int32)
    hip._check(hip.lib().sfh_system_set_segment_grids(s.h, 3, hip._p(seg), hip._p(nx), hip._p(ny)))
    rng = np.random.default_rng(k + 7 * precond)
    kap = np.exp(0.4 * rng.standard_normal(um.nel))
    bcs = [M.dirichlet_all_boundary(m) for m in ms]
    ubc = np.concatenate([b + off[q] for q, b in enumerate(bcs)]).astype(np.int32)
    s.assemble(hip.POISSON, kap, nq1d=2)
    s.applyConstraints(ubc)
    B = rng.standard_normal((um.nnodes, k))
    X = ctx.empty((um.nnodes, k))
    st = hip.SolveStats()
    o = hip.default_opts(tol=1e-12, precond=precond)
    Bd = ctx.to_device(B)
    hip._check(hip.lib().sfh_solve_batched(s.h, 3, hip._p(seg), 0, k, Bd.ptr, X.ptr, C.byref(o), C.byref(st)))
    # Jacobi with at most 4 columns stays on the CSR kernels (4-column tiles)
    assert st.layout == (1 if precond == 1 and k <= 4 else 2) and st.precond == (1 if precond == 1 else 2)
    Xh = X.download().reshape(um.nnodes, k)
    e0 = 0
    for q, m in enumerate(ms):
        kq = kap[e0:e0 + m.nel]
        e0 += m.nel
        rp, ci, va, _ = orc.assemble(m, orc.POISSON, kq, nq1d=2, bc_dofs=bcs[q])
        ref = orc.LU(rp, ci, va).solve(B[off[q]:off[q + 1]])
        got = Xh[off[q]:off[q + 1]]
        assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max(), q


def test_batched_forward_gauss_full_size_vs_fp64_oracle(ctx, orc):
    """The batched-uneven workload at its real sizes (64 problems, n in
    [32, 256], 16-128 sensors): the batched forward model of four problems
    (the smallest and the largest grid, two between) against the FP64 oracle,
    mean and C_u within 1e-8."""
    from oracle.forward import ForwardGauss2D
    from statfem_amd import configs, hip
    from statfem_amd.batch import BatchedForwardGauss
    probs = configs.c4_inputs(seed=0)
    bf = BatchedForwardGauss(ctx, probs)
    means, Cus, st = bf.updateSolution(hip.default_opts(tol=1e-10))
    order = np.argsort([p["n"] for p in probs])
    for q in (order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]):
        p = probs[q]
        ref = ForwardGauss2D(p["mesh"], p["bc"], 1.0).forward_gauss(p["cols"], p["vals"], p["sigma_f"],
                                                                     p["ell_f"], p["reg"])
        em = np.abs(means[q] - ref["mean"]).max() / np.abs(ref["mean"]).max()
        ec = np.abs(Cus[q] - ref["Cu"]).max() / np.abs(ref["Cu"]).max()
        print(f"problem {q}: n {p['n']}, {len(p['cols'])} sensors: mean {em:.1e}, C_u {ec:.1e}", flush=True)
        assert em <= 1e-8 and ec <= 1e-8, (q, em, ec)


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_narrow_column_tiles_match_oracle(ctx, orc, lanes):
    """A batch whose problems carry different column counts (the rest of each
    problem's columns zero) solved with 8-, 16- or 32-column tiles (lanes per
    node capped by sfh_solve_opts.lanes) on the multigrid path, and with
    padding and converged tiles skipped: every problem against its own
    direct solve, padding columns exactly zero."""
    from statfem_amd import hip, mesh as M
    ms = [M.grid2d(40, 33), M.grid2d(17, 17), M.grid2d(52, 29)]
    kq = [5, 23, 40]
    um, off = M.union(ms)
    seg = np.ascontiguousarray(off, np.int32)
    s = hip.SystemSolveHip(ctx, um)
    nx = np.array([len(m.grid[0]) for m in ms], np.int32)
    ny = np.array([len(m.grid[1]) for m in ms], np.int32)
    hip._check(hip.lib().sfh_system_set_segment_grids(s.h, 3, hip._p(seg), hip._p(nx), hip._p(ny)))
    rng = np.random.default_rng(31 + lanes)
    kap = np.exp(0.5 * rng.standard_normal(um.nel))
    bcs = [M.dirichlet_all_boundary(m) for m in ms]
    s.assemble(hip.POISSON, kap, nq1d=2)
    s.applyConstraints(np.concatenate([b + off[q] for q, b in enumerate(bcs)]).astype(np.int32))
    k = max(kq)
    B = rng.standard_normal((um.nnodes, k))
    for q in range(3):
        B[off[q]:off[q + 1], kq[q]:] = 0.0
    X = ctx.empty((um.nnodes, k))
    st = hip.SolveStats()
    o = hip.default_opts(tol=1e-12, lanes=lanes)
    Bd = ctx.to_device(B)
    hip._check(hip.lib().sfh_solve_batched(s.h, 3, hip._p(seg), 1, k, Bd.ptr, X.ptr, C.byref(o), C.byref(st)))
    assert st.precond == 2 and st.layout == 2 and st.converged == 1, st.as_dict()
    Xh = X.download().reshape(um.nnodes, k)
    e0 = 0
    for q, m in enumerate(ms):
        kk = kap[e0:e0 + m.nel]
        e0 += m.nel
        rp, ci, va, _ = orc.assemble(m, orc.POISSON, kk, nq1d=2, bc_dofs=bcs[q])
        ref = orc.LU(rp, ci, va).solve(B[off[q]:off[q + 1]], trans=True)
        got = Xh[off[q]:off[q + 1]]
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max(), q
        assert np.all(got[:, kq[q]:] == 0.0)


def test_autotuned_launch_shape_is_persisted(ctx, tmp_path, monkeypatch):
    """tune_solve probes (lanes, RHS tile) candidates once, stores the winner
    under the device and problem key and returns it from the cache on the
    next call; the solve with the stored shape converges."""
    from statfem_amd import autotune, configs, hip
    from statfem_amd.batch import BatchedForwardGauss
    monkeypatch.setenv("STATFEM_AMD_TUNE_CACHE", str(tmp_path / "shapes.json"))
    probs = configs.c4_inputs(seed=3, nprob=6, nmin=20, nmax=60, smin=4, smax=40)
    bf = BatchedForwardGauss(ctx, probs)
    c1, cost = autotune.tune_solve(bf.sys, bf.Pt, bf.kmax, trans=True, seg=bf.seg, lanes=(4, 16), tiles=(0, 16),
                                   probe_iters=8, repeats=1)
    assert c1["source"] == "probe" and set(cost) == {"4/0", "16/0", "4/16", "16/16"}
    assert cost[f"{c1['lanes']}/{c1['rhs_tile']}"] == min(cost.values())
    c2, cost2 = autotune.tune_solve(bf.sys, bf.Pt, bf.kmax, trans=True, seg=bf.seg, lanes=(4, 16), tiles=(0, 16))
    assert c2["source"] == "cache" and cost2 == {}
    assert (c2["lanes"], c2["rhs_tile"]) == (c1["lanes"], c1["rhs_tile"])
    key = autotune.problem_key(ctx.info(), bf.sys.ndof, bf.kmax, True, bf.seg, 1e-10)
    assert key in autotune.load_cache()
    means, Cus, st = bf.updateSolution(hip.default_opts(tol=1e-10, lanes=c2["lanes"], rhs_tile=c2["rhs_tile"]))
    assert st["converged"] == 1 and all(np.isfinite(Cq).all() for Cq in Cus)


@pytest.mark.parametrize("mg_split", [0, 1])
def test_batched_multigrid_coarse_levels_one_launch_or_per_level(ctx, orc, mg_split):
    """A batch whose largest grid (201 x 201 nodes) starts the single-launch
    coarse end of the V-cycle two levels down while its small grids reach
    their coarsest level earlier (their coarse solve then runs in that
    launch): every problem matches its direct solve, with the coarse levels
    in one launch and with one launch per level."""
    from statfem_amd import hip, mesh as M
    sizes = [(200, 200), (9, 9), (57, 31), (120, 70)]
    meshes = [M.grid2d(a, b) for a, b in sizes]
    mu, off = M.union(meshes)
    s = hip.SystemSolveHip(ctx, mu)
    kap = np.exp(0.5 * np.random.default_rng(8).normal(size=mu.nel))
    s.assemble(hip.POISSON, kap, nq1d=2)
    bc = np.concatenate([M.dirichlet_all_boundary(m) + off[q] for q, m in enumerate(meshes)]).astype(np.int32)
    s.applyConstraints(bc)
    seg = np.ascontiguousarray(off, dtype=np.int32)
    nx = np.array([a + 1 for a, b in sizes], np.int32)
    ny = np.array([b + 1 for a, b in sizes], np.int32)
    hip._check(hip.lib().sfh_system_set_segment_grids(s.h, len(sizes), hip._p(seg), hip._p(nx), hip._p(ny)))
    rp, ci, va, _ = orc.assemble(mu, orc.POISSON, kap, nq1d=2, bc_dofs=bc)
    lu = orc.LU(rp, ci, va)
    rng = np.random.default_rng(12)
    k = 9
    B = rng.normal(size=(mu.nnodes, k))
    B[off[1]:off[2], 4:] = 0.0
    for trans in (0, 1):
        X = ctx.empty((mu.nnodes, k))
        st = hip.SolveStats()
        o = hip.default_opts(tol=1e-11, mg_split=mg_split)
        Bd = ctx.to_device(B)
        hip._check(hip.lib().sfh_solve_batched(s.h, len(sizes), hip._p(seg), trans, k, Bd.ptr, X.ptr, C.byref(o),
                                               C.byref(st)))
        assert st.precond == 2 and st.layout == 2 and st.converged == 1, st.as_dict()
        assert st.inner_iters < 40, st.as_dict()
        Xh = X.download()
        ref = lu.solve(B, trans=bool(trans))
        for q in range(len(sizes)):
            a, b = off[q], off[q + 1]
            assert np.abs(Xh[a:b] - ref[a:b]).max() <= 1e-9 * np.abs(ref[a:b]).max(), (q, trans)
        assert np.all(Xh[off[1]:off[2], 4:] == 0.0)


@pytest.mark.parametrize("nx,ny,k", [(65, 129, 5), (129, 65, 130), (64, 65, 3)])
def test_separable_forcing_covariance_one_row_past_the_tiles(ctx, nx, ny, k):
    """Grid axes of 64 q + 1 nodes: the GEMM tile that also computes the row
    past the last whole tile gives the separable kernel product of the
    definition."""
    from statfem_amd import hip
    rng = np.random.default_rng(nx + ny + k)
    ax, ay = np.sort(rng.uniform(0, 1, nx)), np.sort(rng.uniform(0, 1, ny))
    N = nx * ny
    W = rng.normal(size=(N, k))
    cv = rng.uniform(0.5, 1.5, size=N)
    sig, ell = 0.7, 0.2
    Wd, cvd, V = ctx.to_device(W), ctx.to_device(cv), ctx.empty((N, k))
    hip._check(hip.lib().sfh_covforcing_grid(ctx.h, nx, ny, 1, hip._p(ax), hip._p(ay), None, cvd.ptr, sig, ell, k,
                                             Wd.ptr, V.ptr))
    Kx = np.exp(-(ax[:, None] - ax[None, :]) ** 2 / (2 * ell * ell))
    Ky = np.exp(-(ay[:, None] - ay[None, :]) ** 2 / (2 * ell * ell))
    T = (cv[:, None] * W).reshape(ny, nx, k)
    ref = sig * sig * cv[:, None] * np.einsum("ab,ij,bjc->aic", Ky, Kx, T).reshape(N, k)
    assert np.abs(V.download() - ref).max() <= 1e-12 * np.abs(ref).max()
