"""Benchmark: assembled+solved DOF-RHS/s on one MI355X.

Workload (one step): Poisson on the unit square, 512 x 512 bilinear Q1 cells
(263,169 DOFs), element diffusivity exp(GP(ell=0.1, sigma=0.5)), homogeneous
Dirichlet boundary, 256 sensors. A step is
  GPU assembly (matrix FP64 + FP32 shadow, body force) + row-zero constraints
  + R^T = A^-T P^T           (256-RHS PCG on the declared grid: multigrid
                               V-cycle preconditioner, FP64 recurrences)
  + mean = R F, V = C_f R^T  (separable SE forcing covariance, FP64 MFMA)
  + C_u = R C_f R^T          (Gram, FP64 MFMA)
  + C_u P^T = A^-1 V (+reg)  (second 256-RHS solve) and ubar = A^-1 F.
All inputs are resident in HBM before the timed region. value =
263,169 * 256 / (seconds per step).

Prints one JSON line (rank 0), with the other legs (3D elasticity, batched
problems, surrogate MCMC, conductivity MCMC, SPDE GP, perturbation) as keys of
the same line. --gpus N > 1 is not supported: exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(args):
    """HBM bytes per launch of the C2 step's kernels, measured by this run: two
    rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over a one-step
    child run of this benchmark, each scaled by what the same counter reads
    for a 1 GiB device-to-device copy in that run. Runs before this process
    touches the GPU. Returns {kernel name: bytes per launch} (the largest-grid
    dispatch of each name), or None when the profiler is not there or fails."""
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        log("rocprofv3 not found: traffic from the committed counter summary")
        return None
    from statfem_amd import rocprof_summary as rs
    tmp = tempfile.mkdtemp(prefix="sfh_pmc_", dir="/tmp")
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-cpu", "--no-c1",
             "--no-c5", "--no-c3", "--no-c4", "--no-lhs", "--no-gpq", "--no-fpert", "--copy-cal", "--no-traffic",
             "--n", str(args.n), "--ns", str(args.ns), "--tol", str(args.tol)]
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        log(f"counter pass {counter}")
        cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "--output-format", "csv",
               "-d", os.path.join(tmp, counter), "-o", "c2", "--"] + child
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, env=env, cwd="/tmp")
        if r.returncode != 0:
            log(f"counter pass {counter} failed ({r.returncode}): traffic from the committed counter summary")
            return None
    try:
        t = rs.traffic(os.path.join(tmp, "FETCH_SIZE", "c2_counter_collection.csv"),
                       os.path.join(tmp, "WRITE_SIZE", "c2_counter_collection.csv"))
    except (OSError, KeyError, ValueError) as e:
        log(f"counter summary failed ({e}): traffic from the committed counter summary")
        return None
    out = {}
    for key, v in t["kernels"].items():
        name, grid = key.rsplit(" grid=", 1)
        if name not in out or int(grid) > out[name][0]:
            out[name] = (int(grid), v["bytes"])
    log(f"measured traffic (scales fetch {t['fetch_scale']:.3f}, write {t['write_scale']:.3f})")
    return {k: b for k, (g, b) in out.items()}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (e.g. "k_dir_spmv<float, 16>") from the
    committed counter summary (FETCH_SIZE and WRITE_SIZE passes scaled by a
    copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "pcg_c2_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + " "):
            try:
                return float(line.split()[-1])
            except ValueError:
                return None
    return None


def gpu_run(args, traffic=None):
    from statfem_amd import configs, hip

    n, ns = args.n, args.ns
    cfg = configs.c2_inputs(seed=0, n=n, ns=ns)
    m = cfg["mesh"]
    log(f"mesh {m.nnodes} nodes, {m.nel} elements, {ns} sensors")
    ctx = hip.Context(0)
    s = hip.SystemSolveHip(ctx, m)
    s.set_grid()
    kappa = ctx.to_device(cfg["kappa"])
    fg = hip.ForwardGaussRHSHip(s, cfg["cols"], cfg["vals"], reg_eps=cfg["reg"])
    opts = hip.default_opts(tol=args.tol, rhs_tile=args.tile)
    bc = cfg["bc"]

    def step():
        s.assemble(hip.POISSON, kappa, nq1d=2)
        s.assemble_rhs(nq1d=2, fconst=1.0)
        s.applyConstraints(bc)
        return fg.updateSolution(cfg["sigma_f"], cfg["ell_f"], want_field=True, opts=opts)

    for w in range(args.warmup):
        _, _, st = step()
        log(f"warmup {w}: {st['inner_iters']} inner iterations, {st['outer']} refinements, "
            f"max rel. residual {st['max_relres']:.2e}, {st['ms']:.1f} ms")
    kk = args.tile if 0 < args.tile < ns else ns
    ctx.kernel_sampling(every=args.sample_every, max_samples=4096, min_rhs=kk)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    stats = []
    for k in range(args.steps):
        _, _, st = step()
        stats.append(st)
        log(f"step {k}: {st['inner_iters']} inner iterations, max rel. residual {st['max_relres']:.2e}")
    dev_ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    value = m.nnodes * ns / (wall / args.steps)

    kt = ctx.kernel_times()
    log(f"sampled inner-PCG launches: {kt}")
    # parity sample for the error against the FP64 oracle (after the timed
    # region): the last step's projected mean and R^T on the first columns,
    # from the same solve path as the step
    mean_last = fg.mean.copy()
    P = np.zeros((m.nnodes, ns))
    for j in range(ns):
        P[cfg["cols"][j], j] += cfg["vals"][j]
    Rt_dev, _ = s.solve(P, trans=True, opts=opts)
    kcols = min(ns, args.err_cols)
    rt_sample = Rt_dev.download()[:, :kcols].copy()
    del Rt_dev
    # assembly alone, averaged: the device time of its kernels (HIP events
    # around them, sfh_assembly_times) and of the whole call with its status
    # read-back
    ctx.kernel_sampling(every=args.sample_every, max_samples=4096, min_rhs=kk)
    ctx.timer_start()
    for _ in range(10):
        s.assemble(hip.POISSON, kappa, nq1d=2)
    asm_call_ms = ctx.timer_stop() / 10
    asm_ms, _ = ctx.assembly_times()
    s.applyConstraints(bc)
    kk = args.tile if 0 < args.tile < ns else ns
    N, nnz = s.ndof, s.nnz
    kp = (kk + 3) // 4 * 4
    vec = 4.0 * N * kp                                    # one FP32 N x kp block
    if stats[-1]["layout"] == 2 and stats[-1]["precond"] == 2:
        # multigrid-preconditioned iteration: 9 FP64 stencil coefficients per
        # node; direction: z, p_old read, p written (FP32); update: p read
        # (FP32), FP64 r read and written, and on every other iteration FP64 x
        # read and written with the previous p read (the steps of x go in
        # pairs): 7.5 FP32 words per node and column on average
        algo = {"dir_spmv": 72.0 * N + 3 * vec, "cg_update": 72.0 * N + 7.5 * vec}
        # the survey's FP32 model of the same two phases (per node and column:
        # direction 3 words, update 5 words; FP32 CSR matrix once per 64-column tile)
        mat32 = (8.0 * nnz + 4.0 * (N + 1)) * max(1, kp // 64)
        algo_8d = {"dir_spmv": mat32 + 3 * vec, "cg_update": mat32 + 5 * vec}
        # the update has one instantiation per x-step parity; the timing
        # samples hold as many of each, so its traffic is their mean
        kname = {"dir_spmv": "k_dir_spmv_mg<{lpr}>",
                 "cg_update": ("k_cg_update_mg<{lpr}, true>", "k_cg_update_mg<{lpr}, false>")}
        # V-cycle, all levels: per level FP32 stencil + smoother weight (40 B
        # per node, read twice); level 0 reads the FP64 residual twice and
        # writes z; coarse level l reads r_l twice and writes r_l, e_l; the
        # restricted residual and the coarse correction add a quarter vector
        vc, Nl = 0.0, float(N)
        for lev in range(stats[-1]["levels"]):
            v = 4.0 * Nl * kp
            vc += 80.0 * Nl + (5.0 * v if lev == 0 else 4.0 * v) + 0.5 * v
            Nl = Nl / 4.0
        algo_vc = vc
    elif stats[-1]["layout"] == 2:
        # 9 FP32 stencil coefficients per node (+ D^-1 and diag for the update)
        algo = {"dir_spmv": 36.0 * N + 3 * vec, "cg_update": 44.0 * N + 5 * vec}
        kname = {"dir_spmv": "k_dir_spmv_t2<{lpr}>", "cg_update": "k_cg_update_t2<{lpr}>"}
    else:
        mat = 8.0 * nnz + 4.0 * (N + 1)                   # packed (column, value) entries + row pointers
        algo = {"dir_spmv": mat + 3 * vec, "cg_update": mat + 8.0 * N + 5 * vec}
        kname = {"dir_spmv": "k_dir_spmv<float, {lpr}>", "cg_update": "k_cg_update<float, {lpr}>"}
    if stats[-1]["layout"] != 2 or stats[-1]["precond"] != 2:
        algo_8d = dict(algo)
    per = {k: {"ms": kt[k], "GBs": algo_8d[k] / (kt[k] * 1e-3) / 1e9, "bytes": algo_8d[k],
               "GBs_fp64_recurrences": algo[k] / (kt[k] * 1e-3) / 1e9, "bytes_fp64_recurrences": algo[k],
               "samples": kt["samples"]}
           for k in algo}
    if stats[-1]["precond"] == 2 and kt.get("vcycle", 0) > 0:
        per["vcycle"] = {"ms": kt["vcycle"], "GBs": algo_vc / (kt["vcycle"] * 1e-3) / 1e9, "bytes": algo_vc,
                         "samples": kt["samples"], "launches": 2 * (stats[-1]["levels"] - 1) + 1}
    dom = max(("dir_spmv", "cg_update"), key=lambda k: kt[k])
    lpr = min(16 if stats[-1]["layout"] == 2 else 64, 1 << max(0, (kp // 4 - 1).bit_length()))
    for k in algo:
        names = kname[k] if isinstance(kname[k], tuple) else (kname[k],)
        kns = [n.format(lpr=lpr) for n in names]
        if traffic is not None and all(kn in traffic for kn in kns):
            per[k]["traffic"] = sum(traffic[kn] for kn in kns) / len(kns)
            per[k]["traffic_source"] = "counter passes of this run"
        else:
            got = [profiled_traffic(kn) for kn in kns]
            per[k]["traffic"] = sum(got) / len(got) if all(g is not None for g in got) else None
            per[k]["traffic_source"] = "profiles/pcg_c2_counters.txt"
    asm_bytes = m.nnodes * 2 * 8 + m.nel * 4 * 4 + m.nel * 8 + nnz * 8 + nnz * 4
    dense = dense_stages(ctx, s, ns, cfg)
    res = {
        "ms_per_step": ms_step, "value": value, "dev_ms_per_step": dev_ms / args.steps,
        "inner_iters": stats[-1]["inner_iters"], "outer": stats[-1]["outer"],
        "max_relres": max(st["max_relres"] for st in stats),
        "kernel": dom, "kernels": per, "layout": {1: "csr", 2: "stencil9"}.get(stats[-1]["layout"], "?"),
        "precond": {1: "jacobi", 2: "multigrid"}.get(stats[-1]["precond"], "?"), "mg_levels": stats[-1]["levels"],
        "assembly": {"kernel": "k_quad_grid (closed-form Q1 cell integrals)", "ms": asm_ms,
                     "GBs": asm_bytes / (asm_ms * 1e-3) / 1e9, "bytes": asm_bytes, "call_ms": asm_call_ms},
        "dense": dense,
        "N": N, "nnz": nnz,
        "mean": mean_last, "rt_sample": rt_sample,
    }
    return res, per[dom]


FP64_MFMA_PEAK_TFLOPS = 78.6


def dense_stages(ctx, s, k, cfg, reps=5):
    """FP64-MFMA stages of the step at its sizes, timed alone on the library
    stream: the Gram product C_u = R V (N x k by N x k) and the separable
    forcing covariance V = C_f R^T (one batched GEMM per grid axis)."""
    from statfem_amd import hip
    N = s.ndof
    X = ctx.empty((N, k))
    V = ctx.empty((N, k))
    X.zero()
    V.zero()
    hip.gram(ctx, X, V)
    ctx.timer_start()
    for _ in range(reps):
        hip.gram(ctx, X, V)
    t_gram = ctx.timer_stop() / reps
    cv = s.lumped_mass(2)

    def cf():
        hip._check(hip.lib().sfh_covforcing_apply(s.h, cv.ptr, cfg["sigma_f"], cfg["ell_f"], k, X.ptr, V.ptr, 2))

    cf()
    ctx.timer_start()
    for _ in range(reps):
        cf()
    t_cf = ctx.timer_stop() / reps
    nx, ny = (len(a) for a in s.mesh.grid)
    f_gram = 2.0 * N * k * k
    f_cf = 2.0 * k * nx * ny * (nx + ny)
    out = {}
    for name, f, t in (("gram", f_gram, t_gram), ("covforcing_separable", f_cf, t_cf)):
        tf = f / (t * 1e-3) / 1e12
        out[name] = {"ms": t, "gflop": f / 1e9, "tflops": tf, "peak_tflops": FP64_MFMA_PEAK_TFLOPS,
                     "frac": tf / FP64_MFMA_PEAK_TFLOPS, "bound": "mfma"}
    log(f"dense stages: {out}")
    return out


def c5_run(args):
    """Hyperparameter MCMC with the learned surrogate (256 x 256 Q1 Poisson,
    200 sensors, 8 snapshots, 64 chains). The statFEM prior at the sensors
    comes from the GPU forward model; a tuned pilot sets the random-walk step
    sizes shared by both samplers; the plain chain runs the reference's
    random-walk Metropolis on the exact likelihood; the surrogate run pays for
    its training design, its training and its sampling. ESS/s is the minimum
    over the three hyperparameters."""
    from statfem_amd import configs, hip, inference as I
    cfg = configs.c5_inputs(seed=0, n=args.c5_n)
    m = cfg["mesh"]
    ctx = hip.Context(0)
    s = hip.SystemSolveHip(ctx, m)
    s.set_grid()
    s.assemble(hip.POISSON, 1.0, nq1d=2)
    s.assemble_rhs(nq1d=2, fconst=1.0)
    s.applyConstraints(cfg["bc"])
    fg = hip.ForwardGaussRHSHip(s, cfg["cols"], cfg["vals"], reg_eps=cfg["reg"])
    t0 = time.perf_counter()
    mean, Cu, _ = fg.updateSolution(cfg["sigma_f"], cfg["ell_f"], opts=hip.default_opts(tol=1e-10))
    t_prior = time.perf_counter() - t0
    ns, no = len(cfg["xs"]), cfg["no"]
    Y = hip.LikelihoodKOHHip(ctx, cfg["xs"], Cu, mean, np.zeros((no, ns)), cfg["sigma_e"]).giveTypicalReadings(
        *cfg["truth"], no, seed=cfg["data_seed"])
    lik = hip.LikelihoodKOHHip(ctx, cfg["xs"], Cu, mean, Y, cfg["sigma_e"])
    prior = dict(prior_mode=2, prior_mu=cfg["prior_mu"], prior_var=cfg["prior_var"])
    rng = np.random.default_rng(3)
    chains = args.c5_chains
    th0 = cfg["prior_mu"] + 0.1 * rng.standard_normal((chains, 3))
    prop, start, center, scale, t_pilot = I.tuned_proposal(lik, th0, steps=2000, **prior)
    log(f"C5 prior {t_prior:.2f}s, pilot {t_pilot:.2f}s, proposal sd {prop}")
    t0 = time.perf_counter()
    plain, acc = lik.mcmc(start, prop, args.c5_steps, seed=1, **prior)
    t_plain = time.perf_counter() - t0
    ess_p, rate_p = I.ess_per_second(plain, t_plain)
    log(f"C5 plain: {t_plain:.2f}s, acceptance {acc.mean():.3f}, ESS {ess_p.round(0)}")
    da = I.surrogate_posterior(lik, start, prop, center, scale, args.c5_da_steps, args.c5_sub, seed=5, **prior)
    t_da = sum(da["times"].values())
    ess_d, rate_d = I.ess_per_second(da["samples"], t_da)
    log(f"C5 surrogate: {da['times']}, acceptance {da['acc'].mean():.3f}, ESS {ess_d.round(0)}")
    burn_p, burn_d = int(0.1 * plain.shape[1]), int(0.1 * da["samples"].shape[1])
    em, es = I.moment_errors(da["samples"][:, burn_d:], plain[:, burn_p:])
    return {
        "workload": f"Poisson {args.c5_n}x{args.c5_n} Q1, {ns} sensors, {no} snapshots, {chains} chains",
        "plain": {"steps": args.c5_steps, "s": t_plain, "acceptance": float(acc.mean()),
                  "ess": ess_p.tolist(), "ess_per_s": rate_p},
        "surrogate": {"steps": args.c5_da_steps, "sub_steps": args.c5_sub, "s": t_da,
                      "phases_s": da["times"], "acceptance": float(da["acc"].mean()),
                      "sub_acceptance": float(da["sub_acc"].mean()), "train_points": da["n_train"],
                      "train_loss": da["train_loss"], "ess": ess_d.tolist(), "ess_per_s": rate_d},
        "ess_per_s_ratio": rate_d / rate_p,
        "moment_rel_err": {"mean": em.tolist(), "sd": es.tolist()},
        "prior_setup_s": t_prior, "pilot_s": t_pilot,
    }


def c1_run(args):
    """Config C1: 64 x 64 cells split into P1 triangles (4,225 nodes), f = 1 +
    GP(0.1, 0.2), 32 interior sensors, 10 readings. One statFEM update on the
    GPU: re-assembly and constraints, the 32-column transposed solve, the
    separable forcing covariance and the Gram product (prior mean and C_u at
    the sensors, tolerance 1e-12), then the log-likelihood of 64 (rho,
    sigma_d, ell_d) candidates; 20 timed updates after one untimed one. CPU:
    the oracle's update (sparse direct LU, the dense forcing-covariance
    projection, the same 64 likelihoods) on one core."""
    from statfem_amd import configs, hip, mesh as M
    c = configs.c1_inputs(0)
    m = c["mesh"]
    ctx = hip.Context(0)
    s = hip.SystemSolveHip(ctx, m)
    rng = np.random.default_rng(3)
    thetas = np.column_stack([rng.uniform(0.7, 1.3, 64), rng.uniform(0.01, 0.1, 64), rng.uniform(0.1, 0.5, 64)])
    opts = hip.default_opts(tol=1e-12)
    state = {}

    def update():
        s.assemble(hip.POISSON, 1.0, nq1d=1)
        s.assemble_rhs(nq1d=1, fconst=1.0)
        s.applyConstraints(c["bc"])
        if "fg" not in state:
            s.set_grid()
            state["fg"] = hip.ForwardGaussRHSHip(s, c["cols"], c["vals"], reg_eps=c["reg"])
        mean, Cu, st = state["fg"].updateSolution(c["sigma_f"], c["ell_f"], opts=opts)
        Y = c["rho"] * mean[None, :] + c["d"] + c["e"]
        lik = hip.LikelihoodKOHHip(ctx, c["xs"], Cu, mean, Y, c["sigma_e"])
        ll, stt = lik.giveLogLikelihood(thetas)
        if not np.all(stt == 0):
            raise RuntimeError("C1 likelihood failed")
        return st, Y

    update()
    ctx.sync()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        st, Y = update()
    ctx.sync()
    t = (time.perf_counter() - t0) / reps
    log(f"C1 {m.nnodes} nodes, {len(c['cols'])} sensors: {1e3 * t:.2f} ms per update with 64 likelihoods "
        f"({st['inner_iters']} inner iterations, precond {st['precond']})")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        P = M.dense_projection(m.nnodes, c["cols"], c["vals"])
        t0 = time.perf_counter()
        F1 = orc.body_force(m, nq1d=1, fconst=1.0)
        rp, ci, va, F1 = orc.assemble(m, orc.POISSON, 1.0, nq1d=1, bc_dofs=c["bc"], F=F1)
        R = orc.LU(rp, ci, va).solve(P.T.copy(), trans=True).T.copy()
        cv = orc.body_force(m, nq1d=1, fconst=1.0)
        cv[c["bc"]] = 0.0
        Cu = orc.covforcing_project(m.coords, cv, c["sigma_f"], c["ell_f"], R) \
            + (c["sigma_f"] * c["reg"]) ** 2 * P @ P.T
        mean = R @ F1
        for th in thetas:
            orc.loglik(Cu, mean, Y, orc.mismatch_cov(c["xs"], th[1], th[2], c["sigma_e"]), th[0])
        tc = time.perf_counter() - t0
        cpu = {"value": 1.0 / tc, "unit": "updates/s", "cores": 1, "kind": "port",
               "sample": f"one update: oracle assembly + sparse LU + 32 transposed solves + dense C_f projection "
                         f"+ 64 likelihoods, {tc:.2f}s"}
    return {
        "workload": f"statFEM update on 64x64 P1 ({m.nnodes} nodes), 32 sensors x 10 readings, "
                    f"64 likelihood candidates",
        "ms_per_update": 1e3 * t, "updates_per_s": 1.0 / t, "inner_iters": st["inner_iters"],
        "precond": st["precond"], "cpu_baseline": cpu, "gpu_vs_cpu": (1.0 / t) / cpu["value"] if cpu else None,
    }


def c3_run(args):
    """3D linear elasticity with an SPDE Matern body force: n^3 hex8 cells
    (64^3: 823,875 DOFs), 128 top-face sensors. Prior at the sensors (one
    128-RHS transposed elasticity solve, one 128-RHS Matern solve, a weighted
    Gram), a synthetic reading from a seeded Matern load, and the posterior:
    sensor mean and variance, log-determinant, and the Woodbury field mean
    (one more Matern and one more elasticity solve). One untimed pass first;
    the rate counts the assembly and the timed prior."""
    from statfem_amd import configs, hip
    from statfem_amd.spde_elasticity import SPDEElasticity
    cfg = configs.c3_inputs(seed=0, n=args.c3_n)
    ctx = hip.Context(0)
    opts = hip.default_opts(tol=1e-10)
    pipe = SPDEElasticity(ctx, cfg, opts=opts, opts_matern=hip.default_opts(tol=1e-10))
    # one untimed pass (first-use allocations, solver work areas and the 3D
    # hierarchy), then the timed prior and posterior
    mean, Cu = pipe.prior()
    y, g, u = pipe.synthesize()
    pipe.posterior(mean, Cu, y, field=True)
    ctx.sync()
    mean, Cu = pipe.prior()
    t_prior = pipe.times["assemble"] + pipe.times["prior"]
    post = pipe.posterior(mean, Cu, y, field=True)
    st = pipe.stats
    log(f"C3 {pipe.ndof} DOFs: assembly {pipe.times['assemble']:.2f}s, prior {pipe.times['prior']:.2f}s "
        f"(R^T {st['Rt']['inner_iters']} it, Matern {st['matern']['inner_iters']} it), "
        f"posterior {pipe.times['posterior']:.2f}s")
    fm = post["field_mean"]
    # elasticity assembly alone (hex8 cells of the declared grid: closed-form
    # cell integrals summed per (node, neighbour) block into the CSR values,
    # FP64 + FP32 copy): device time of the kernels per call (HIP events), and
    # of the whole call with its status read-back
    se = pipe.se
    se.assemble(hip.ELASTICITY, cfg["E"], cfg["nu"], nq1d=cfg["nq1d"])
    ctx.kernel_sampling(every=15, max_samples=64)
    ctx.timer_start()
    for _ in range(5):
        se.assemble(hip.ELASTICITY, cfg["E"], cfg["nu"], nq1d=cfg["nq1d"])
    asm_call_ms = ctx.timer_stop() / 5
    asm_ms, _ = ctx.assembly_times()
    ctx.kernel_sampling(every=0)
    se.applyConstraints(cfg["bc"])
    m = cfg["mesh"]
    asm_bytes = 12.0 * se.nnz + 4.0 * 3 * m.nnodes                # CSR values FP64 + FP32 written, row starts read
    asm_flop = 64 * 80.0 * m.nnodes                               # 64 (cell, corner) blocks per node, ~80 flops each
    bound_ms = max(asm_bytes / (HBM_PEAK_GBS * 1e9), asm_flop / (FP64_PEAK_TFLOPS * 1e12)) * 1e3
    log(f"C3 elasticity assembly {asm_ms:.3f} ms (bound {bound_ms:.3f} ms)")
    # the same pipeline at the size the FP64 oracle covers (16^3), for the error
    small = None
    if not args.no_cpu:
        cfg16 = configs.c3_inputs(seed=0, n=16)
        p16 = SPDEElasticity(ctx, cfg16, opts=opts, opts_matern=hip.default_opts(tol=1e-10))
        m16, C16 = p16.prior()
        y16, _, _ = p16.synthesize()
        post16 = p16.posterior(m16, C16, y16, field=True)
        small = {"cfg": cfg16, "mean": m16, "Cu": C16, "y": y16, "post": post16, "field": p16._field.download()}
    return {
        "workload": f"elasticity {args.c3_n}^3 hex8, {pipe.ndof} DOFs, {pipe.ns} sensors, SPDE Matern load",
        "dofs": pipe.ndof, "rhs": pipe.ns,
        "assembled_solved_dof_rhs_per_s": pipe.ndof * pipe.ns / t_prior,
        "s": {"assemble": pipe.times["assemble"], "prior": pipe.times["prior"], "posterior": pipe.times["posterior"]},
        "iterations": {k: v["inner_iters"] for k, v in st.items()},
        "max_relres": max(v["max_relres"] for v in st.values()),
        "logdet": post["logdet"], "posterior_sd_mean": float(np.sqrt(post["var"]).mean()),
        "field_mean_vs_truth_rel": float(np.abs(fm - u).max() / np.abs(u).max()),
        "assembly_roofline": {"kernel": "k_hex_grid<elasticity>", "ms": asm_ms, "bytes": asm_bytes, "gflop": asm_flop / 1e9,
                              "bound": "fp64" if asm_flop / (FP64_PEAK_TFLOPS * 1e12) > asm_bytes / (HBM_PEAK_GBS * 1e9)
                              else "hbm",
                              "bound_ms": bound_ms, "frac": bound_ms / asm_ms, "call_ms": asm_call_ms},
    }, small


def c4_run(args):
    """64 independent Poisson problems of uneven size (n ~ U{32..256}, Q1,
    ell_f ~ U(0.05, 0.3), 16-128 sensors). Batched: one block-diagonal system,
    one segmented multi-RHS solve with per-problem scalars, tile width from the
    autotuner; against the same priors computed one problem at a time (systems
    assembled beforehand on both sides; the timed part is the forward-model
    update: solves, C_f, Gram); error per problem against an FP64 solve
    (precision 1, tol 1e-12)."""
    from statfem_amd import configs, hip
    from statfem_amd.autotune import tune_solve
    from statfem_amd.batch import BatchedForwardGauss, SequentialForwardGauss
    probs = configs.c4_inputs(seed=0, nprob=args.c4_nprob)
    ctx = hip.Context(0)
    dofs_rhs = float(sum(p["mesh"].nnodes * len(p["cols"]) for p in probs))
    t0 = time.perf_counter()
    bf = BatchedForwardGauss(ctx, probs)
    t_setup = time.perf_counter() - t0
    choice, cost = tune_solve(bf.sys, bf.Pt, bf.kmax, trans=True, seg=bf.seg, tiles=(0, 64))
    log(f"C4 autotuner: {choice} (device ms of the capped probe solve per lanes/rhs_tile: {cost})")
    opts = hip.default_opts(tol=1e-10, rhs_tile=choice["rhs_tile"], lanes=choice["lanes"])
    bf.updateSolution(opts)                      # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    means, Cus, st = bf.updateSolution(opts)
    t_batch = time.perf_counter() - t0
    seq = SequentialForwardGauss(ctx, probs)
    seq.updateSolution(hip.default_opts(tol=1e-10, rhs_tile=0))              # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    sm, sC, sit = seq.updateSolution(hip.default_opts(tol=1e-10, rhs_tile=0))
    t_seq = time.perf_counter() - t0
    rm, rC, _ = seq.updateSolution(hip.default_opts(tol=1e-12, precision=1))
    seq.close()
    # (against the library's own FP64-recurrence solve of every problem; the
    # error against the FP64 oracle is computed in the CPU leg on a sample)
    err = [max(np.abs(C - R).max() / np.abs(R).max(), np.abs(m - r).max() / np.abs(r).max())
           for C, R, m, r in zip(Cus, rC, means, rm)]
    log(f"C4 batched {t_batch:.3f}s ({st['inner_iters']} it), sequential {t_seq:.3f}s ({sit} it), "
        f"max rel diff vs the GPU FP64 path {max(err):.2e}")
    return {
        "workload": f"{len(probs)} Poisson Q1 problems, n in [32, 256], {int(dofs_rhs)} DOF-RHS in total",
        "dofs": int(bf.mesh.nnodes), "launch_shape": choice, "probe_ms": cost,
        "batched_s": t_batch, "sequential_s": t_seq, "speedup": t_seq / t_batch, "setup_s": t_setup,
        "batched_dof_rhs_per_s": dofs_rhs / t_batch, "sequential_dof_rhs_per_s": dofs_rhs / t_seq,
        "batched_iters": st["inner_iters"], "sequential_iters": sit,
        "max_rel_diff_vs_gpu_fp64": float(max(err)),
    }, {"probs": probs, "means": means, "Cus": Cus}


def lhs_run(args):
    """Conductivity (LHS) uncertainty with every proposal re-solved
    (ForwardGaussLHS + the laplaceInverse MCMC, re-solving): `lhs_chains` chains
    on one n x n Q1 Poisson problem whose conductivity is the GP mean of 5
    anchor values; each MCMC step interpolates, assembles and solves every
    chain's field in one batched pass (one grid copy and multigrid hierarchy
    per chain) and evaluates the 64-sensor likelihood. CPU sample: the oracle's
    assembly and banded direct solve of a few proposals (1 thread)."""
    from statfem_amd import configs, hip
    from statfem_amd.lhs import ForwardGaussLHSHip
    cfg = configs.lhs_inputs(seed=0, n=args.lhs_n)
    m, C, steps = cfg["mesh"], args.lhs_chains, args.lhs_steps
    ctx = hip.Context(0)
    opts = hip.default_opts(tol=1e-10)
    t0 = time.perf_counter()
    fg = ForwardGaussLHSHip(ctx, m, cfg["bc"], cfg["anchors"], cfg["gp"], cfg["cols"], cfg["vals"], copies=C,
                            nq1d=2, opts=opts)
    t_setup = time.perf_counter() - t0
    # synthetic readings: the forward model at the true anchor values plus noise
    mean, _ = fg.updateSolution(np.r_[np.log(cfg["w_true"]), np.log(cfg["sigma_e"])])
    rng = np.random.default_rng(cfg["noise_seed"])
    Y = mean[0][:, None] + cfg["sigma_e"] * rng.standard_normal((len(cfg["cols"]), cfg["no"]))
    fg.set_data(Y, cfg["xs"], rho=1.0, mis_sigma=0.0, mis_ell=0.3, prior_mu=cfg["prior_mu"],
                prior_var=cfg["prior_var"])
    th0 = np.repeat(cfg["init"][None, :], C, axis=0)
    fg.mcmc(th0, cfg["propsd"], 2, seed=3, resolve=True, keep_samples=False)   # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    samples, acc, st = fg.mcmc(th0, cfg["propsd"], steps, seed=1, resolve=True)
    t = time.perf_counter() - t0
    N = m.nnodes
    props = C * steps / t
    log(f"LHS {C} chains x {steps} re-solving steps on {N}-DOF grids: {t:.2f}s ({props:.0f} proposals/s), "
        f"{st['inner_iters']} inner iterations, acceptance {acc.mean():.2f}, precond {st['precond']}")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        k, t0 = 0, time.perf_counter()
        while k < 2 or (time.perf_counter() - t0 < 10.0 and k < 50):
            orc.lhs_solution(m, cfg["bc"], cfg["anchors"], np.exp(samples[k % C, k % steps, :5]), *cfg["gp"])
            k += 1
        tc = time.perf_counter() - t0
        cpu = {"value": k / tc, "unit": "proposals/s", "cores": 1, "kind": "port",
               "sample": f"{k} proposals: oracle assembly + banded LU solve of the {N}-DOF system each"}
    fg.close()
    return {
        "workload": f"conductivity MCMC, {C} chains, Poisson {args.lhs_n}x{args.lhs_n} Q1 ({N} DOFs per chain), "
                    f"GP conductivity from 5 anchors, {len(cfg['cols'])} sensors x {cfg['no']} readings, "
                    f"every proposal re-assembled and re-solved",
        "chains": C, "steps": steps, "s": t, "setup_s": t_setup, "proposals_per_s": props,
        "dof_solves_per_s": props * N, "inner_iters": st["inner_iters"], "precond": st["precond"],
        "acceptance": float(acc.mean()), "cpu_baseline": cpu,
        "gpu_vs_cpu": props / cpu["value"] if cpu else None,
    }


def gpq_run(args):
    """Sparse-precision SPDE GP likelihood surface (GaussianProcessQ through
    DriverMaxLikelihood, nu = 1, beta = 1): a 256 x 256 Q1 grid (66,049
    nodes), 200 interior sensors, 8 readings drawn from the model; 8 length
    scales x 16 noise levels = 128 log-likelihood evaluations in one call
    (the noise levels of one length scale share its 200-column solve). The
    same surface at nu = 0.5 (beta = 0.75, fractional: the degree-4 rational
    factors add 4 sparse products and 4 more 200-column solves per length
    scale) is timed beside it. CPU sample: the same sensor-space computation
    for one length scale with the oracle's banded LU of A (1 thread)."""
    from statfem_amd import hip, mesh as M
    from statfem_amd.gpq import GaussianProcessQHip
    n, ns, no = args.gpq_n, 200, 8
    m = M.grid2d(n)
    rng = np.random.default_rng(21)
    xs = rng.uniform(0.02, 0.98, size=(ns, 2))
    cols, vals = M.sensor_rows(m, xs)
    Y = np.sin(3 * xs[:, :1]) * np.cos(2 * xs[:, 1:]) + 0.05 * rng.standard_normal((ns, no))
    ctx = hip.Context(0)
    s = hip.SystemSolveHip(ctx, m)
    s.set_grid()
    gp = GaussianProcessQHip(s, cols, vals, Y, nq1d=2)
    nu, sigma = 1.0, 1.0
    ells = np.geomspace(0.05, 0.6, 8)
    epss = np.geomspace(0.01, 0.5, 16)
    sets = [[sigma, np.sqrt(2 * nu) / ell, nu, e] for ell in ells for e in epss]
    opts = hip.default_opts(tol=1e-10)
    gp.giveLogLikelihood(sets[:16], opts)     # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    ll, st, stats = gp.giveLogLikelihood(sets, opts)
    t = time.perf_counter() - t0
    best = int(np.argmax(ll))
    log(f"GPQ {m.nnodes} nodes, {ns} sensors x {no}: {len(sets)} log-likelihoods in {t:.3f}s "
        f"({stats['inner_iters']} inner iterations, precond {stats['precond']}); best ell "
        f"{ells[best // 16]:.3f}, eps {epss[best % 16]:.3f}")
    nuf = 0.5
    fsets = [[sigma, np.sqrt(2 * nuf) / ell, nuf, e] for ell in ells for e in epss]
    gp.giveLogLikelihood(fsets[:16], opts)    # warm-up (the rational factors are computed here)
    ctx.sync()
    t0 = time.perf_counter()
    llf, stf, statsf = gp.giveLogLikelihood(fsets, opts)
    tf = time.perf_counter() - t0
    if not np.all(stf == 0):
        raise RuntimeError("fractional SPDE GP likelihood failed")
    log(f"GPQ nu = {nuf} (beta = 0.75, rational degree 4): {len(fsets)} log-likelihoods in {tf:.3f}s "
        f"({statsf['inner_iters']} inner iterations)")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        kap = np.sqrt(2 * nu) / ells[0]
        rp, ci, va, _ = orc.assemble(m, orc.MATERN, kap, 1.0, nq1d=2, factor=1.0 / kap ** 2)
        mass = orc.body_force(m, nq1d=2, fconst=1.0)
        P = M.dense_projection(m.nnodes, cols, vals)
        X = orc.LU(rp, ci, va).solve(P.T.copy())
        G = X.T @ (mass[:, None] * X) / (orc.matern_tau(sigma, kap, nu, 2) ** 2 * kap ** 4)
        for e in epss:
            B = e * e * np.eye(ns) + no * G
            np.linalg.slogdet(B)
            np.linalg.solve(B, Y.sum(axis=1))
        tc = time.perf_counter() - t0
        cpu = {"value": len(epss) / tc, "unit": "log-likelihoods/s", "cores": 1, "kind": "port",
               "sample": f"one length scale (16 noise levels): oracle assembly + banded LU + {ns} solves on "
                         f"{m.nnodes} nodes + sensor-space algebra, {tc:.1f}s"}
    gp.close()
    return {
        "workload": f"SPDE GP (nu = 1) on {n}x{n} Q1 ({m.nnodes} nodes), {ns} sensors x {no} readings, "
                    f"{len(sets)} log-likelihoods (8 length scales x 16 noise levels)",
        "s": t, "loglik_per_s": len(sets) / t, "inner_iters": stats["inner_iters"], "precond": stats["precond"],
        "fractional": {"nu": nuf, "beta": 0.75, "degree": 4, "s": tf, "loglik_per_s": len(fsets) / tf,
                       "inner_iters": statsf["inner_iters"]},
        "cpu_baseline": cpu, "gpu_vs_cpu": (len(sets) / t) / cpu["value"] if cpu else None,
    }


def fpert_run(args):
    """First-order perturbation (ForwardPerturbation) with an element-wise
    conductivity: Poisson on an n x n Q1 grid with homogeneous Dirichlet,
    one material parameter per element (n^2 parameters, their stiffness
    derivatives and centroids), 32 sensors, SE forcing covariance; one
    updateSolution (u, A^-1 by one N-column solve, F = A^-1 C_f A^-T + u u^T,
    the structured product over parameter pairs, R C_f R^T). CPU sample: the
    oracle on a 16 x 16 grid (256 parameters), with the GPU timed there too."""
    from statfem_amd import hip, mesh as M
    from statfem_amd.fpert import ForwardPerturbationHip, conductivity_derivatives
    prm = {"forcing.sigma": 0.3, "forcing.length": 0.2, "material.sigma": 0.2, "material.length": 0.25}
    ctx = hip.Context(0)
    opts = hip.default_opts(tol=1e-10)

    def build(n):
        m = M.grid2d(n)
        bc = M.dirichlet_all_boundary(m)
        kap = np.exp(0.3 * np.sin(7 * m.coords[m.conn].mean(axis=1)[:, 0]))
        s = hip.SystemSolveHip(ctx, m)
        s.set_grid()
        s.assemble(hip.POISSON, kap, nq1d=2)
        s.assemble_rhs(nq1d=2, fconst=1.0)
        s.applyConstraints(bc)
        rng = np.random.default_rng(3)
        xs = rng.uniform(0.05, 0.95, size=(32, 2))
        cols, vals = M.sensor_rows(m, xs)
        dofs, blocks = conductivity_derivatives(m, bc)
        fp = ForwardPerturbationHip(s, cols, vals, dofs, blocks, m.coords[m.conn].mean(axis=1))
        return m, bc, kap, s, cols, vals, dofs, blocks, fp

    def timed(fp, reps):
        fp.updateSolution(prm, opts)              # warm-up
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fp.updateSolution(prm, opts)
        return (time.perf_counter() - t0) / reps

    m, bc, kap, s, cols, vals, dofs, blocks, fp = build(args.fpert_n)
    t = timed(fp, 3)
    fp.close()
    ms, bcs, kaps, ss, cs, vs, ds, bs, fps = build(16)
    ts = timed(fps, 5)
    fps.close()
    log(f"FPERT {m.nnodes} nodes, {m.nel} parameters: {1e3 * t:.1f} ms per update; 16^2: {1e3 * ts:.2f} ms")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        F = orc.body_force(ms, nq1d=2, fconst=1.0)
        rp, ci, va, F = orc.assemble(ms, orc.POISSON, kaps, nq1d=2, bc_dofs=bcs, F=F)
        cv = orc.body_force(ms, nq1d=2, fconst=1.0)
        cv[bcs] = 0.0
        P = M.dense_projection(ms.nnodes, cs, vs)
        t0 = time.perf_counter()
        orc.fpert_update(rp, ci, va, F, ms.coords, cv, prm["forcing.sigma"], prm["forcing.length"], P, ds, bs,
                         ms.coords[ms.conn].mean(axis=1), prm["material.sigma"], prm["material.length"])
        tc = time.perf_counter() - t0
        cpu = {"value": 1.0 / tc, "unit": "updates/s", "cores": 1, "kind": "port",
               "sample": f"oracle updateSolution on 16^2 Q1 (289 nodes, 256 parameters): {tc:.2f}s; the GPU at the "
                         f"same size: {1e3 * ts:.2f} ms"}
    return {
        "workload": f"ForwardPerturbation on {args.fpert_n}x{args.fpert_n} Q1 ({m.nnodes} nodes), {m.nel} element "
                    f"conductivities, 32 sensors",
        "ms_per_update": 1e3 * t, "updates_per_s": 1.0 / t, "ms_per_update_16": 1e3 * ts, "cpu_baseline": cpu,
        "gpu_vs_cpu_16": (1.0 / ts) / cpu["value"] if cpu else None,
    }


def cpu_baseline_c3(small):
    """The elasticity statFEM at 16^3 (14,739 DOFs, 128 sensors) by the FP64
    oracle (banded Cholesky of both operators, one host thread for the
    factorisations): the error of the GPU's prior and posterior at that size,
    and the oracle's own time as the CPU baseline of this leg."""
    from oracle.forward import elasticity_spde_reference
    cfg = small["cfg"]
    t0 = time.perf_counter()
    ref = elasticity_spde_reference(cfg, small["y"])
    t_cpu = time.perf_counter() - t0
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    post = small["post"]
    err = {"prior_mean": rel(small["mean"], ref["mean"]), "Cu": rel(small["Cu"], ref["Cu"]),
           "field": rel(small["field"], ref["field"]), "post_mean": rel(post["mean"], ref["post_mean"]),
           "post_var": rel(post["var"], ref["post_var"]),
           "logdet": float(abs(post["logdet"] - ref["logdet"]) / abs(ref["logdet"])),
           "field_mean": rel(post["field_mean"], ref["field_mean"]), "size": "16^3 hex8, 14739 DOFs, 128 sensors"}
    N = 3 * cfg["mesh"].nnodes
    cpu = {"value": N * len(cfg["cols"]) / t_cpu, "unit": "DOF-RHS/s", "cores": 16, "kind": "port",
           "sample": f"oracle elasticity statFEM at 16^3 ({N} DOFs, 128 sensors): assembly, banded Cholesky of the "
                     f"elasticity and Matern operators (1 thread), multi-RHS solves (16 threads), prior, field "
                     f"block and posterior: {t_cpu:.2f}s"}
    log(f"C3 16^3 error vs FP64 oracle: {err}")
    return err, cpu


def cpu_baseline_c4(gpu, nsample=8):
    """Error of the batched forward model against the FP64 oracle on a sample
    of the 64 problems spread over the grid sizes (mean and C_u per problem)."""
    from oracle.forward import ForwardGauss2D
    probs = gpu["probs"]
    order = np.argsort([p["n"] for p in probs])
    pick = sorted(set(int(order[i]) for i in np.linspace(0, len(order) - 1, nsample).round().astype(int)))
    errs = []
    t0 = time.perf_counter()
    for q in pick:
        p = probs[q]
        ref = ForwardGauss2D(p["mesh"], p["bc"], 1.0).forward_gauss(p["cols"], p["vals"], p["sigma_f"], p["ell_f"],
                                                                     p["reg"])
        errs.append(max(float(np.abs(gpu["means"][q] - ref["mean"]).max() / np.abs(ref["mean"]).max()),
                        float(np.abs(gpu["Cus"][q] - ref["Cu"]).max() / np.abs(ref["Cu"]).max())))
    log(f"C4 error vs FP64 oracle on problems {pick}: {errs} ({time.perf_counter() - t0:.1f}s)")
    return {"max": max(errs), "median": float(np.median(errs)), "problems": pick,
            "grid_sizes": [int(probs[q]["n"]) for q in pick],
            "oracle": "banded Cholesky + separable C_f checked by direct sums, FP64"}


def c2_errors(orc, cfg, rp, ci, va, ch, F, res):
    """Relative error of the GPU's C2 outputs against the oracle's FP64 direct
    solves: R^T = A^-T P^T on the sampled columns (2-norm per column, the
    constrained components restored as r_c = p_c - sum_f A[f][c] r_f) and the
    projected mean P A^-1 F on all sensors (max norm)."""
    import scipy.sparse as sp
    from statfem_amd import mesh as M
    m, bc = cfg["mesh"], cfg["bc"]
    N = m.nnodes
    A = sp.csr_matrix((va, ci, rp), shape=(N, N))
    P = M.dense_projection(N, cfg["cols"], cfg["vals"]).T.copy()
    kc = res["rt_sample"].shape[1]
    Pk = P[:, :kc].copy()
    Pf = Pk.copy()
    Pf[bc] = 0.0
    nt = max(1, min(16, len(os.sched_getaffinity(0))))
    R = ch.solve(Pf, threads=nt)
    R[bc] = 0.0
    R[bc] = Pk[bc] - (A.T @ R)[bc]
    e_rt = np.linalg.norm(res["rt_sample"] - R, axis=0) / np.linalg.norm(R, axis=0)
    mref = P.T @ ch.solve(F)
    e_mean = float(np.abs(res["mean"] - mref).max() / np.abs(mref).max())
    return {"Rt": float(e_rt.max()), "mean": e_mean, "Rt_columns": kc, "sensors": int(P.shape[1]),
            "oracle": "banded Cholesky of the eliminated operator, FP64"}


def cpu_baseline(args, budget_s, res=None):
    """The oracle (reference algorithm restated in C++, FP64, 1 thread):
    Triplet sort-merge assembly, banded Cholesky of the eliminated operator,
    then transposed + forward solve pairs for a sample of right-hand sides,
    extrapolated to all 256. The dense C_f product is not included (it would
    be 6.9e10 kernel evaluations per RHS set in the reference form). With the
    GPU results `res` of the same workload it also returns their error against
    these FP64 solves."""
    from oracle import oracle as orc
    from statfem_amd import configs, mesh as M
    cfg = configs.c2_inputs(seed=0, n=args.cpu_n, ns=args.ns)
    m = cfg["mesh"]
    t0 = time.perf_counter()
    F = orc.body_force(m, nq1d=2, fconst=1.0)
    rp, ci, va, F = orc.assemble(m, orc.POISSON, cfg["kappa"], nq1d=2, bc_dofs=cfg["bc"], F=F)
    t_asm = time.perf_counter() - t0
    mask = np.zeros(m.nnodes, np.uint8)
    mask[cfg["bc"]] = 1
    t0 = time.perf_counter()
    ch = orc.Chol(rp, ci, va, mask)
    t_fac = time.perf_counter() - t0
    P = M.dense_projection(m.nnodes, cfg["cols"], cfg["vals"]).T.copy()
    P[cfg["bc"]] = 0.0
    k = max(1, min(args.ns, args.cpu_rhs))
    t0 = time.perf_counter()
    X = ch.solve(P[:, :k])
    ch.solve(X)
    t_solve_k = time.perf_counter() - t0
    t_total = t_asm + t_fac + t_solve_k * args.ns / k
    N = m.nnodes
    one = N * args.ns / t_total
    # the same solves with the right-hand sides split over the host's cores
    # (assembly and factorisation stay on one thread, as in the reference)
    nt = max(1, min(16, len(os.sched_getaffinity(0))))
    kt = max(1, min(args.ns, 2 * nt))
    t0 = time.perf_counter()
    X = ch.solve(P[:, :kt], threads=nt)
    ch.solve(X, threads=nt)
    t_solve_kt = time.perf_counter() - t0
    t_total_mt = t_asm + t_fac + t_solve_kt * args.ns / kt
    err = None
    if res is not None and args.cpu_n == args.n:
        err = c2_errors(orc, cfg, rp, ci, va, ch, F, res)
        log(f"C2 error vs FP64 oracle: {err}")
    return err, {
        "value": N * args.ns / t_total_mt, "unit": "DOF-RHS/s", "cores": nt, "kind": "port",
        "sample": (f"{args.cpu_n}x{args.cpu_n} Q1 ({N} DOFs): assembly {t_asm:.2f}s + banded Cholesky "
                   f"{t_fac:.2f}s (1 thread) + {kt} of {args.ns} solve pairs on {nt} threads {t_solve_kt:.2f}s, "
                   f"extrapolated to {args.ns} RHS; C_f product excluded. One thread throughout: {one:.3g} "
                   f"DOF-RHS/s ({k} solve pairs {t_solve_k:.2f}s)"),
        "single_core": one,
        "cpu_model": cpu_model(), "host_threads": len(os.sched_getaffinity(0)),
    }


def cpu_model():
    """The host CPU's model name (the first processor's entry of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ns", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--sample-every", type=int, default=15,
                    help="time the inner PCG iteration pairs 0, 1 modulo n with HIP events inside the timed solves "
                         "(a pair holds one update with the x step and one without)")
    ap.add_argument("--tile", type=int, default=0, help="right-hand sides per solve (0: all in one solve)")
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--cpu-rhs", type=int, default=8)
    ap.add_argument("--err-cols", type=int, default=32,
                    help="R^T columns compared with the oracle's FP64 solves (with the CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the small statFEM update leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the surrogate-MCMC leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 3D elasticity leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched uneven-problems leg")
    ap.add_argument("--no-lhs", action="store_true", help="skip the conductivity-MCMC leg")
    ap.add_argument("--no-gpq", action="store_true", help="skip the sparse-precision GP leg")
    ap.add_argument("--gpq-n", type=int, default=256)
    ap.add_argument("--no-fpert", action="store_true", help="skip the first-order perturbation leg")
    ap.add_argument("--fpert-n", type=int, default=64)
    ap.add_argument("--lhs-n", type=int, default=128)
    ap.add_argument("--lhs-chains", type=int, default=64)
    ap.add_argument("--lhs-steps", type=int, default=100)
    ap.add_argument("--c4-nprob", type=int, default=64)
    ap.add_argument("--c3-n", type=int, default=64)
    ap.add_argument("--c5-n", type=int, default=256)
    ap.add_argument("--c5-chains", type=int, default=64)
    ap.add_argument("--c5-steps", type=int, default=20000)
    ap.add_argument("--c5-da-steps", type=int, default=2000)
    ap.add_argument("--c5-sub", type=int, default=32)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the kernels' HBM traffic")
    ap.add_argument("--copy-cal", action="store_true",
                    help="also run a 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only (there is no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    traffic = None if args.no_traffic else measured_traffic(args)
    res, dom = gpu_run(args, traffic)
    if args.copy_cal:
        from statfem_amd import hip
        c = hip.Context(0)
        a, b = c.empty((1 << 27,)), c.empty((1 << 27,))
        a.zero()
        import ctypes
        hip._check(hip.lib().sfh_memcpy_d2d(c.h, b.ptr, a.ptr, ctypes.c_size_t(1 << 30)))
        c.sync()
        log("copy calibration: 1 GiB device-to-device")
    c1 = None if args.no_c1 else c1_run(args)
    c5 = None if args.no_c5 else c5_run(args)
    c3, c3_small = (None, None) if args.no_c3 else c3_run(args)
    c4, c4_gpu = (None, None) if args.no_c4 else c4_run(args)
    lhs = None if args.no_lhs else lhs_run(args)
    gpq = None if args.no_gpq else gpq_run(args)
    fpert = None if args.no_fpert else fpert_run(args)
    err, cpu = (None, None) if args.no_cpu else cpu_baseline(args, 30.0, res)
    if c3 is not None and c3_small is not None:
        c3["max_rel_err_vs_fp64"], c3["cpu_baseline"] = cpu_baseline_c3(c3_small)
    if c4 is not None and not args.no_cpu:
        c4["max_rel_err_vs_fp64"] = cpu_baseline_c4(c4_gpu)
    line = {
        "metric": "assembled+solved DOF-RHS/s; MCMC ESS/s surrogate vs plain GPU at 1e-8 rel err",
        "value": res["value"],
        "unit": "DOF-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"Poisson {args.n}x{args.n} Q1, {res['N']} DOFs, exp(GP) diffusivity, "
                               f"{args.ns} sensors/RHS: assembly + A^-T P^T + C_f + Gram + A^-1 C_f R^T",
                   "dofs": res["N"], "rhs": args.ns, "nnz": res["nnz"], "tol": args.tol, "rhs_tile": args.tile,
                   "inner_iters": res["inner_iters"], "max_relres": res["max_relres"],
                   "max_rel_err_vs_fp64": err, "operator_layout": res["layout"],
                   "precond": res["precond"], "mg_levels": res["mg_levels"],
                   "kernels": res["kernels"], "assembly": res["assembly"]},
        "roofline": {"kernel": res["kernel"], "bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["GBs"] / HBM_PEAK_GBS, "traffic": dom.get("traffic"),
                     "traffic_source": dom.get("traffic_source"),
                     "model": "FP32 words per node and column (x, r, p read; x, r written) + FP32 CSR per tile",
                     "achieved_fp64_recurrences": dom["GBs_fp64_recurrences"],
                     "frac_fp64_recurrences": dom["GBs_fp64_recurrences"] / HBM_PEAK_GBS,
                     "bytes_per_launch": dom["bytes"], "bytes_per_launch_fp64_recurrences": dom["bytes_fp64_recurrences"]},
        "cpu_baseline": cpu,
        "statfem_update": c1,
        "mcmc": c5,
        "elasticity": c3,
        "batched": c4,
        "lhs_mcmc": lhs,
        "spde_gp": gpq,
        "perturbation": fpert,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
