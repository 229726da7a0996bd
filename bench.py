"""Benchmark of the fiber3d GMRES-matvec hot path on one MI355X.

Workload (one "step" = one pass of the dominant hot-path kernel over one
batch): the raw Oseen source-target sum with N = M = 65536 points uniform in
the unit cube, forces N(0,1), viscosity 1, self term excluded, evaluated with
the FP32-compensated kernel (FP64 output).  Inputs are resident in HBM before
the timed region.  value = Gpair-interactions/s over the K timed steps.

Also reported on the same line: the FP64 kernel's rate and its error, the
stresslet rates, the roofline of the dominant kernel (FP32 VALU, 28 flops per
Oseen pair) and a CPU baseline (the oracle's C restatement of the reference's
numba loop, OpenMP over targets, on a bounded sample of targets).

"fiber_system" (default: the large workload, 4096 fibers x 64 nodes, a
32768-node periphery and 8 rigid spheres) reports time-steps/s to the 1e-8
FP64 GMRES residual for: plain FP64 (the reference method on the GPU), fast
(FP32-compensated inner matvec, FP32 MFMA fiber LU), extrapolated_fast,
corrected_fast (extrapolation of the residual-corrected history) and
learned_fast (the trained per-node net), all with initial guesses from the
solver's own past steps.  "fiber_system_series" does the same for the time
series (1000 fibers x 32 nodes, 8000-node periphery; the learned predictor
trains on its first 100 steps, then 300 steps are timed).  Each mode runs
its own trajectory from the same seeded state; untimed warm-up steps,
including the predictor's training, precede the timed ones.  Both legs
carry the oracle's CPU time per plain step, extrapolated from samples.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--system 0|3|4|-1]
       [--no-series] [--no-lu] [--no-small] [--no-cpu] [--no-graph]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "time-steps/s at 1e-8 GMRES residual, learned+FP32 vs plain FP64; Gpairs/s"
N_PTS = 65536
OSEEN_FLOPS = 28      # per pair: 3 sub, r^2 5, rsqrt 1, d.f 5, r^-2 1, scale 1, 3 x 2 FMA
STRESSLET_FLOPS = 30
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak (packed)
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
# HBM bytes per launch of the FP32-compensated Oseen pair kernel at N = M = 65536:
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over `bench.py --steps 2
# --warmup 1 --no-cpu --no-lu --no-small --no-series --system -1` (28313.3 and
# 12288 KiB per dispatch), scaled by what the same counters read for
# `bench.py --copy-probe 512` (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  The writes
# are the partial sums (32 source chunks, 4 combined per workgroup) that the
# reduce kernel reads back; the fetches are mostly source records re-read by
# the 1024 target waves.
OSEEN_F32C_HBM_BYTES = 70568576


def host_cores():
    """Host threads this process may use: the CPU affinity set, capped by the
    cgroup CPU quota where one is set; with the CPU model and OpenMP
    settings, for the baseline's report."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cores = min(n, quota) if quota else n
    return {"threads": int(cores), "affinity_cpus": int(n), "cgroup_quota_cpus": quota,
            "cpu_model": model, "OMP_PROC_BIND": "close"}


def _omp(cores):
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ["OMP_PROC_BIND"] = "close"


def cpu_baseline(x, f, n_targets=4096):
    """Oracle C restatement of kernels.py:794-846 on the host (port), on all
    the host threads this process may use."""
    from oracle import build as obuild
    from oracle import cbaseline
    if not os.path.exists(obuild.OUT):
        obuild.build(verbose=False)
    host = host_cores()
    _omp(host["threads"])
    rows = np.arange(n_targets)
    cbaseline.oseen(x, x[:64], f)  # warm-up (thread pool)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        cbaseline.oseen(x, x[rows], f)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return {"value": round(N_PTS * n_targets / dt / 1e9, 4), "unit": "Gpairs/s",
            "cores": host["threads"], "kind": "port", "host": host,
            "sample": f"Oseen FP64, all {N_PTS} sources x {n_targets} targets, median of 5 "
                      "after one warm-up"}


def oracle_step_estimate(cfg, fibers, xs, ns, bodies, iters, sample_targets=2048,
                         sample_fibers=64):
    """CPU time of one plain FP64 step of the reference's method at this
    workload, extrapolated from bounded samples of the oracle on all host
    threads: the O(N^2) pair sums of one matvec (C restatement, OpenMP over
    targets, timed on sample_targets targets and scaled to all targets),
    per-fiber assembly and LAPACK LU (timed on sample_fibers fibers, scaled
    to all), and the per-fiber LU solves of the preconditioner; one step =
    assembly + LU + (iters + 2) x (matvec + preconditioner), iters being the
    plain GPU solve's iteration count (the same Krylov method from zero)."""
    import scipy.linalg as sla
    from oracle import build as obuild
    from oracle import cbaseline as C
    from oracle import fiber as OF
    if not os.path.exists(obuild.OUT):
        obuild.build(verbose=False)
    host = host_cores()
    _omp(host["threads"])
    spec = cfg.fibers
    X = np.concatenate(fibers)
    npts, nper = X.shape[0], xs.shape[0]
    bn = np.concatenate([b[1] for b in bodies]) if bodies else np.zeros((0, 3))
    bnn = np.concatenate([b[2] for b in bodies]) if bodies else np.zeros((0, 3))
    rng = np.random.default_rng(0)
    f = rng.normal(size=X.shape)
    q = rng.normal(size=xs.shape)
    T = X[rng.choice(npts, min(sample_targets, npts), replace=False)]
    C.oseen(X, T[:16], f)  # warm-up
    t0 = time.perf_counter()
    C.oseen(X, T, f)
    t_os = (time.perf_counter() - t0) * (npts + nper + bn.shape[0]) / T.shape[0]
    t0 = time.perf_counter()
    C.stresslet(xs, T, ns, q)
    t_st = (time.perf_counter() - t0) * (npts + nper + bn.shape[0]) / T.shape[0]
    t_bd = 0.0
    if bn.shape[0]:
        t0 = time.perf_counter()
        C.stresslet(bn, T, bnn, rng.normal(size=bn.shape))
        t_bd = (time.perf_counter() - t0) * (npts + nper + bn.shape[0]) / T.shape[0]
    matvec = t_os + t_st + t_bd
    ks = rng.choice(len(fibers), min(sample_fibers, len(fibers)), replace=False)
    c0, c1 = OF.constants(spec.epsilon, cfg.eta)
    t_as = t_lu = t_sv = 0.0
    for k in ks:
        x = fibers[k]
        fm = OF.FiberMatrices(x.shape[0], cfg.nfd)
        t0 = time.perf_counter()
        A = OF.form_linear_operator(fm, x, spec.length, spec.E, cfg.dt, c0, c1, cfg.penalty)
        R = OF.compute_RHS(fm, x, spec.length, cfg.dt, c0, c1, cfg.penalty)
        A_sq, b = OF.apply_BC_rectangular(fm, x, spec.length, spec.E, A, R)
        t1 = time.perf_counter()
        lu = sla.lu_factor(A_sq, check_finite=False)
        t2 = time.perf_counter()
        sla.lu_solve(lu, b, check_finite=False)
        t3 = time.perf_counter()
        t_as += t1 - t0
        t_lu += t2 - t1
        t_sv += t3 - t2
    scale = len(fibers) / len(ks)
    t_as, t_lu, t_sv = t_as * scale, t_lu * scale, t_sv * scale
    step = t_as + t_lu + (iters + 2) * (matvec + t_sv)
    return {"value": round(1.0 / step, 6), "unit": "time-steps/s", "cores": host["threads"],
            "kind": "port", "host": host, "extrapolated": True,
            "seconds_per_step": round(step, 3),
            "parts_s": {"pair_sums_per_matvec": round(matvec, 4),
                        "assembly": round(t_as, 4), "fiber_lu": round(t_lu, 4),
                        "fiber_lu_solves_per_apply": round(t_sv, 4), "matvecs": iters + 2},
            "sample": f"oracle on {T.shape[0]} of {npts + nper + bn.shape[0]} targets (pair "
                      f"sums) and {len(ks)} of {len(fibers)} fibers (assembly, LAPACK LU and "
                      f"solve), scaled linearly; {iters} GMRES iterations (plain GPU solve)"}


def system_leg(cfg_id, steps, train_steps=0, scale=None, shell_pc="auto", with_bodies=True,
               graph=True, cpu=True):
    """time-steps/s of the implicit fiber step, plain FP64 vs fast (+ learned guess).

    Each mode runs its own trajectory from the same seeded initial state:
    `train_steps` untimed steps (the learned mode trains its predictor on them),
    then `steps` timed steps, each solved to the configuration's GMRES
    tolerance with the FP64 residual deciding convergence.  The fast modes
    close each solve with the FP64-anchored residual update and record every
    timed step; after timing, each recorded step is re-assembled and its true
    residual recomputed in FP64 (max_true_residual)."""
    import torch
    from fiber3d_amd import configs
    from fiber3d_amd.learned import LearnedGuess
    from fiber3d_amd.system import Shell
    from fiber3d_amd.tstep import TimeStepper

    cfg = configs.config(cfg_id)
    spec = cfg.fibers
    nper = cfg.periphery_nodes
    if scale is not None:
        spec = configs.FiberSpec(scale[0], spec.num_points, spec.length, spec.ball_radius,
                                 spec.E, spec.epsilon, spec.min_sep, spec.end_force)
        nper = scale[1]
    fibers = configs.make_fibers(spec, cfg.seed)
    xs, ns, ws = configs.make_periphery(nper, cfg.periphery_radius)
    shell = Shell(xs, ns, ws, eta=cfg.eta, pc=shell_pc, pc_f32=True)
    bodies = None
    if cfg_id == 4 and with_bodies:
        # 8 free rigid spheres, radius 0.5, 600 nodes, no overlap (margin 0.1)
        bodies = configs.make_bodies(8, 0.5, 600, spec.ball_radius, fibers=fibers, seed=cfg.seed)
    out = {"workload": f"{cfg.name}: {len(fibers)} fibers x {spec.num_points} nodes, "
                       f"{nper}-node periphery"
                       + (f", {len(bodies)} rigid spheres x 600 nodes" if bodies else "")
                       + f", dt={cfg.dt}, GMRES tol {cfg.gmres_tol}",
           "shell_preconditioner": shell.pc, "arnoldi_operator": "hip graph" if graph else "eager"}
    modes = [("plain", "plain", None), ("fast", "fast", None)]
    if train_steps:
        modes.append(("extrapolated_fast", "fast", LearnedGuess(train_steps=10**9, use_net=False)))
        modes.append(("corrected_fast", "fast",
                      LearnedGuess(train_steps=10**9, use_net=False, correct=True)))
        modes.append(("learned_fast", "fast", LearnedGuess(train_steps=train_steps, epochs=20,
                                                          max_samples=32768)))
    learned_warm = max([p.warm_steps for _, _, p in modes if p is not None] or [1])
    for name, mode, pred in modes:
        ts = TimeStepper(fibers, shell, mode=mode, eta=cfg.eta, dt=cfg.dt, E=spec.E,
                         epsilon=spec.epsilon, tol=cfg.gmres_tol, restart=cfg.restart,
                         maxiter=cfg.maxiter, predictor=pred, lengths=np.full(len(fibers), spec.length),
                         bodies=bodies, graph=graph, record=mode == "fast")
        # a predictor needs `keep` solutions before its full guess; the learned
        # mode then picks its baseline (4 steps), collects train_steps
        # samples, trains and checks the net on 4 further steps, all before
        # timing starts.  Every guessing mode times the same steps: those
        # after the learned mode's window (held-out steps).
        warm = 1 if pred is None else learned_warm
        for _ in range(warm):
            ts.step()
        ts.records = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [ts.step() for _ in range(steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        its = [st.iterations for st in stats]
        g0 = [st.guess_residual for st in stats if st.guess_residual is not None]
        out[name] = {"time_steps_per_s": round(steps / dt, 4), "mean_gmres_iters": float(np.mean(its)),
                     "zero_iteration_steps": int(sum(i == 0 for i in its)),
                     "mean_guess_residual": float(np.mean(g0)) if g0 else None,
                     "max_true_residual": float(max(st.residual for st in stats)),
                     "steps": steps, "untimed_steps": warm, "lu_refactor_every": ts.lu_every}
        if ts.record:
            true, drhs = ts.verify()
            kinds = [st.residual_kind for st in stats]
            out[name].update({
                "max_true_residual": float(max(true)),
                "max_residual_bound": float(max(st.residual_bound for st in stats)),
                "closing_residual": {k: kinds.count(k) for k in sorted(set(kinds))},
                "inner_rel_err": ts.inner_rel_err,
                "verify_rhs_max_diff": drhs})
        ts.records = []
        if pred is not None and pred.use_net:
            out[name]["predictor"] = pred.report
    if cpu and "plain" in out:
        out["cpu_baseline"] = oracle_step_estimate(
            cfg, fibers, xs, ns, bodies or [], int(round(out["plain"]["mean_gmres_iters"])))
    if "plain" in out and "fast" in out:
        out["speedup_fast_vs_plain"] = round(out["fast"]["time_steps_per_s"] / out["plain"]["time_steps_per_s"], 3)
    for m in ("extrapolated_fast", "corrected_fast", "learned_fast"):
        if m in out:
            out[f"speedup_{m}_vs_plain"] = round(
                out[m]["time_steps_per_s"] / out["plain"]["time_steps_per_s"], 3)
    return out


def small_leg(steps=30, cpu=True):
    """The small workload (32 fibers x 32 nodes, 1024-node periphery): GPU
    time-steps/s, plain FP64 and fast with the extrapolated guess, beside one
    step of the oracle's FP64 restatement on the host (NumPy/SciPy: per-step
    assembly, per-fiber LU, GMRES; the one-time shell matrix and its LU are
    set up once outside the timed step, as on the GPU)."""
    import torch
    from fiber3d_amd import configs
    from fiber3d_amd.learned import LearnedGuess
    from fiber3d_amd.system import Shell
    from fiber3d_amd.tstep import TimeStepper

    cfg = configs.config(0)
    spec = cfg.fibers
    fibers = configs.make_fibers(spec, cfg.seed)
    xs, ns, ws = configs.make_periphery(cfg.periphery_nodes, cfg.periphery_radius)
    shell = Shell(xs, ns, ws, eta=cfg.eta)
    out = {"workload": f"{cfg.name}: {len(fibers)} fibers x {spec.num_points} nodes, "
                       f"{cfg.periphery_nodes}-node periphery, dt={cfg.dt}, GMRES tol {cfg.gmres_tol}"}
    for name, mode, pred in (("plain", "plain", None),
                             ("extrapolated_fast", "fast",
                              LearnedGuess(train_steps=10**9, use_net=False))):
        ts = TimeStepper(fibers, shell, mode=mode, eta=cfg.eta, dt=cfg.dt, E=spec.E,
                         epsilon=spec.epsilon, tol=cfg.gmres_tol, restart=cfg.restart,
                         maxiter=cfg.maxiter, predictor=pred,
                         lengths=np.full(len(fibers), spec.length))
        for _ in range(4 if pred is None else pred.keep + 2):
            ts.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [ts.step() for _ in range(steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"time_steps_per_s": round(steps / dt, 2),
                     "mean_gmres_iters": float(np.mean([st.iterations for st in stats])),
                     "max_residual": float(max(st.residual for st in stats))}
    if cpu:
        from oracle import tstep as OT
        host = host_cores()
        kw = dict(eta=cfg.eta, dt=cfg.dt, E=spec.E, epsilon=spec.epsilon,
                  lengths=np.full(len(fibers), spec.length))
        # one-time set-up (shell subtraction vectors, dense shell matrix and
        # its LU: tstep.py:278-356) outside the timed step, as on the GPU
        t0 = time.perf_counter()
        first = OT.HydroSystem(fibers, (xs, ns, ws), **kw)
        t_init = time.perf_counter() - t0
        cache = first.shell_state()
        times, its, res = [], 0, 0.0
        for _ in range(3):
            t0 = time.perf_counter()
            ora = OT.HydroSystem(fibers, (xs, ns, ws), shell_cache=cache, **kw)
            xo, info, its = OT.gmres(ora.matvec, ora.rhs, ora.precond, tol=cfg.gmres_tol,
                                     restart=cfg.restart, maxiter=cfg.maxiter)
            times.append(time.perf_counter() - t0)
            res = float(np.linalg.norm(ora.rhs - ora.matvec(xo)) / np.linalg.norm(ora.rhs))
        step = float(np.median(times))
        out["cpu_baseline"] = {"value": round(1.0 / step, 4), "unit": "time-steps/s",
                               "cores": host["threads"], "kind": "port", "host": host,
                               "sample": "one plain FP64 step of this workload by the oracle "
                                         "(NumPy/SciPy restatement of the reference's step: "
                                         "per-step assembly, fiber LU, GMRES), median of 3; "
                                         "the one-time shell set-up excluded",
                               "one_time_setup_s": round(t_init, 3),
                               "gmres_iters": int(its), "residual": res}
    return out


def lu_leg(reps=5, cpu=True):
    """Batched fiber preconditioner at workload 2: 2048 fibers, N in
    {16,24,32,48,64} (block order n = 4N), A_sq from the fiber discretisation.
    One persistent factor launch (FP64, or FP32 + FP64 refinement) and one
    solve launch for all fibers."""
    import torch
    from fiber3d_amd import configs
    from fiber3d_amd.dense import BatchedLU
    from fiber3d_amd.system import FiberSystem

    cfg = configs.config(2)
    fibers = configs.make_fibers(cfg.fibers, cfg.seed)
    S = FiberSystem(fibers, None, lu_precision="f64")
    n = S.bs.sizes.astype(np.float64)
    flops = float(np.sum(2.0 / 3.0 * n**3))
    rng = np.random.default_rng(2)
    xs = torch.from_numpy(rng.normal(size=S.bs.nrows)).cuda()
    b = torch.empty_like(xs)
    S.gA(S.A_buf, xs, b)
    nn = float(np.sum(n**2))
    out = {"workload": f"{len(fibers)} fibers, N in {cfg.fibers.num_points}, n = 4N",
           "factor_flops": flops}
    # A_sq x for every block (block_gemv_kernel): HBM-bound, 8 n^2 bytes per block
    S.gA(S.A_buf, xs, b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        S.gA(S.A_buf, xs, b)
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / reps
    out["block_gemv"] = {"ms": round(tg * 1e3, 4), "bytes": 8 * nn,
                         "roofline": {"bound": "hbm", "achieved": round(8 * nn / tg / 1e9, 1),
                                      "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                      "frac": round(8 * nn / tg / 1e9 / PEAK_HBM_GBS, 3)}}
    for prec in ("f64", "f32c"):
        lu = BatchedLU(S.bs, S.A_buf, precision=prec, refine=10, tol=1e-12)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            lu.factor()
        torch.cuda.synchronize()
        tf = (time.perf_counter() - t0) / reps
        y = lu.solve(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            y = lu.solve(b)
        torch.cuda.synchronize()
        ts = (time.perf_counter() - t0) / reps
        err = float(torch.linalg.norm(y - xs) / torch.linalg.norm(xs))
        sweeps = int(lu.iters.max()) if prec == "f32c" else 0
        # the time stepper's solve: no refinement (the preconditioner)
        yp = lu.solve(b, refine=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            lu.solve(b, out=yp, refine=0)
        torch.cuda.synchronize()
        tp = (time.perf_counter() - t0) / reps
        # solve bytes: the factor once per substitution, plus A (FP64) per
        # refinement sweep; sweeps counted at the block maximum
        fb = 8 if prec == "f64" else 4
        sbytes = nn * (fb * (1 + sweeps) + 8 * sweeps)
        peak = PEAK_FP64_TFLOPS if prec == "f64" else PEAK_FP32_TFLOPS
        out[prec] = {"factor_ms": round(tf * 1e3, 3), "factor_gflops": round(flops / tf / 1e9, 1),
                     "factor_roofline": {"bound": "valu" if prec == "f64" else "mfma",
                                         "achieved": round(flops / tf / 1e12, 3), "peak": peak,
                                         "unit": "TFLOP/s",
                                         "frac": round(flops / tf / 1e12 / peak, 4)},
                     "solve_ms": round(ts * 1e3, 3), "forward_error": err,
                     "max_refine_sweeps": sweeps,
                     "solve_roofline": {"bound": "hbm", "achieved": round(sbytes / ts / 1e9, 1),
                                        "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                        "frac": round(sbytes / ts / 1e9 / PEAK_HBM_GBS, 3)},
                     "plain_solve_ms": round(tp * 1e3, 4),
                     "plain_solve_roofline": {"bound": "hbm",
                                              "achieved": round(nn * fb / tp / 1e9, 1),
                                              "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                              "frac": round(nn * fb / tp / 1e9 / PEAK_HBM_GBS,
                                                            3)}}
    if cpu:
        import scipy.linalg as sla
        blocks = S.bs.unpack(S.A_buf)[:256]
        t0 = time.perf_counter()
        for A in blocks:
            sla.lu_factor(A, check_finite=False)
        dt = time.perf_counter() - t0
        nb = np.array([A.shape[0] for A in blocks], float)
        out["cpu_baseline"] = {"kind": "library (SciPy/LAPACK dgetrf per fiber, host)",
                               "sample": "first 256 of the fibers", "factor_gflops":
                               round(float(np.sum(2 / 3 * nb**3)) / dt / 1e9, 2)}
    return out


def mlp_leg(reps=20):
    """The learned guess's hidden layer at the large workload: one BF16 MFMA
    GEMM with bias + ReLU epilogue over every fiber node, (262144 x 256) x
    (256 x 256), FP32 accumulate (f3d_gemm_bf16)."""
    import torch
    from fiber3d_amd import learned
    M, K, N = 262144, 256, 256
    g = torch.Generator().manual_seed(0)
    A = torch.randn(M, K, generator=g).cuda()
    B = (torch.randn(K, N, generator=g) / 16).cuda()
    bias = torch.randn(N, generator=g).cuda()
    C = torch.empty(M, N, device="cuda")
    learned.gemm(A, B, C, epi=learned.EPI_BIAS_RELU, bias=bias)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        learned.gemm(A, B, C, epi=learned.EPI_BIAS_RELU, bias=bias)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ref = torch.relu(A.to(torch.bfloat16).float() @ B.to(torch.bfloat16).float() + bias)
    err = float(torch.linalg.norm(C - ref) / torch.linalg.norm(ref))
    fl = 2.0 * M * N * K
    # FP32 operands in and out: 63 flop/byte at this shape, below the BF16 MFMA
    # ridge (2500 TFLOP/s / 8 TB/s = 312 flop/byte), so HBM bounds the GEMM
    nbytes = 4.0 * (M * K + K * N + M * N + N)
    return {"shape": [M, K, N], "ms": round(dt * 1e3, 4), "rel_err_vs_fp32_of_bf16_inputs": err,
            "tflops": round(fl / dt / 1e12, 1),
            "roofline": {"bound": "hbm", "achieved": round(nbytes / dt / 1e9, 1),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(nbytes / dt / 1e9 / PEAK_HBM_GBS, 4),
                         "flop_per_byte": round(fl / nbytes, 1)}}


def copy_probe(mib):
    """Three device copies of `mib` MiB (the calibration of the PMC byte
    counters), then 20 more timed with HIP events: the read + write rate of
    a plain streaming copy, the practical ceiling of kernels that read one
    array and write another of the same size."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * a.numel() * 4
    print(json.dumps({"copy_probe_bytes_per_launch": {"read": a.numel() * 4,
                                                      "write": a.numel() * 4},
                      "launches": 3 + reps, "timed_ms": round(ms, 4),
                      "read_write_gbs": round(nbytes / ms / 1e6, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--system", type=int, default=4,
                    help="also time implicit steps of fiber-system workload 0, 3 or 4 "
                         "(-1: raw sweep only)")
    ap.add_argument("--system-steps", type=int, default=50)
    ap.add_argument("--train-steps", type=int, default=30,
                    help="untimed steps the learned predictor trains on (0: plain and fast only)")
    ap.add_argument("--no-series", dest="series", action="store_false",
                    help="skip the time series (workload 3: 1000 fibers x 32 nodes, 8000-node "
                         "periphery; the learned predictor trains on its first steps)")
    ap.add_argument("--series-steps", type=int, default=300)
    ap.add_argument("--series-train", type=int, default=100)
    ap.add_argument("--no-lu", dest="lu", action="store_false",
                    help="skip the batched fiber LU (workload 2) and MLP GEMM legs")
    ap.add_argument("--no-small", dest="small", action="store_false",
                    help="skip the small-workload time steps (GPU and oracle on the host)")
    ap.add_argument("--shell-pc", default="auto", choices=["auto", "dense", "block"])
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="run the Arnoldi steps eagerly instead of replaying recorded HIP graphs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 plain device copies of MIB MiB (calibrates the HBM "
                         "byte counters FETCH_SIZE / WRITE_SIZE against a known size)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "sharding over several GPUs is designed but not measured here.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fiber3d_amd import kernels as FK

    dev = FK.device()
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 1.0, size=(N_PTS, 3))
    f = rng.normal(size=(N_PTS, 3))
    nrm = rng.normal(size=(N_PTS, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    mu = rng.normal(size=(N_PTS, 3))
    wq = rng.uniform(0.5, 1.5, size=N_PTS) / N_PTS
    X = torch.from_numpy(x).to(dev)
    F = torch.from_numpy(f).to(dev)
    NRM = torch.from_numpy(nrm).to(dev)
    Q = torch.from_numpy(mu * wq[:, None]).to(dev)
    out = torch.empty((N_PTS, 3), dtype=torch.float64, device=dev)
    pairs = float(N_PTS) * N_PTS

    def run(kernel, precision, steps, warmup, nrm_t=None, den=F):
        for _ in range(warmup):
            FK.pairsum(kernel, X, den, X, src_nrm=nrm_t, exclude_self=True, precision=precision,
                       out=out)
        torch.cuda.synchronize()
        # HIP events around the pair kernel alone, recorded by libf3d on the
        # stream it launches on (the caller's current stream)
        FK.kernel_timer(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            FK.pairsum(kernel, X, den, X, src_nrm=nrm_t, exclude_self=True, precision=precision,
                       out=out)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms, launches = FK.kernel_timer_read()
        FK.kernel_timer(False)
        assert launches == steps, (launches, steps)
        return wall / steps * 1e3, kms / launches, out.clone()

    # headline: FP32-compensated Oseen
    ms32, k32, u32 = run("oseen", "f32c", args.steps, args.warmup)
    ms64, k64, u64 = run("oseen", "f64", max(3, args.steps // 4), 1)
    err32 = float(torch.linalg.norm(u32 - u64) / torch.linalg.norm(u64))
    sms32, sk32, s32 = run("stresslet", "f32c", max(3, args.steps // 4), 1, nrm_t=NRM, den=Q)
    sms64, sk64, s64 = run("stresslet", "f64", max(3, args.steps // 4), 1, nrm_t=NRM, den=Q)
    serr = float(torch.linalg.norm(s32 - s64) / torch.linalg.norm(s64))

    gp = pairs / (ms32 / 1e3) / 1e9
    achieved = OSEEN_FLOPS * pairs / (k32 / 1e3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(gp, 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (compensated; FP64 output)",
        "data": "synthetic (seeded: 65536 points U[0,1)^3, forces N(0,1))",
        "config": {"workload": "raw Oseen source-target sum, N=M=65536, self term excluded, "
                               "FP32-compensated", "n_sources": N_PTS, "n_targets": N_PTS,
                   "eta": 1.0},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": OSEEN_F32C_HBM_BYTES, "traffic_unit": "bytes/launch",
                     "kernel": "pairsum_kernel<float,OSEEN>", "flops_per_pair": OSEEN_FLOPS,
                     "kernel_ms": round(k32, 4)},
        "oseen_f32c_rel_err_vs_f64": err32,
        "oseen_f64_gpairs_s": round(pairs / (ms64 / 1e3) / 1e9, 3),
        "oseen_f64_tflops": round(OSEEN_FLOPS * pairs / (k64 / 1e3) / 1e12, 3),
        "stresslet_f32c_gpairs_s": round(pairs / (sms32 / 1e3) / 1e9, 3),
        "stresslet_f64_gpairs_s": round(pairs / (sms64 / 1e3) / 1e9, 3),
        "stresslet_f32c_rel_err_vs_f64": serr,
    }
    if args.lu:
        line["batched_lu"] = lu_leg(cpu=not args.no_cpu)
        line["mlp_gemm_bf16"] = mlp_leg()
    if args.small:
        line["fiber_system_small"] = small_leg(cpu=not args.no_cpu)
    if args.series:
        line["fiber_system_series"] = system_leg(3, args.series_steps, args.series_train,
                                                 shell_pc=args.shell_pc, graph=args.graph,
                                                 cpu=not args.no_cpu)
    if args.system >= 0:
        line["fiber_system"] = system_leg(args.system, args.system_steps, args.train_steps,
                                          shell_pc=args.shell_pc, graph=args.graph,
                                          cpu=not args.no_cpu)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(x, f)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
