"""Autotuner of the launch shapes: pair sums per size bucket, the batched LU
and the BF16 GEMM.

Pair sums: for each (kernel, precision) and each tuned problem size
(n_targets, n_sources) of a grid of powers of two, it times f3d_pairsum over
the launch shapes -- targets per lane K in {1, 2, 4} and the wavefront count
the source split aims for -- and keeps the fastest.  apply() installs the
fastest shape of the nearest tuned size (in log2 n_targets + log2 n_sources)
into every size bucket of the library (f3d_pairsum_set_launch_bucket), and
the 65536 x 65536 shape as the (kernel, precision) default.  The large
fiber -> fiber Oseen sum (262144 nodes, exclusion ranges) is tuned on its
own layout, bucket 18/18.

Batched LU (f3d_lu_set_launch): waves per CU of the one-wave factorisation,
and LDS-staged versus streaming solves, timed on fiber blocks of the shapes
the stepper factors.  BF16 GEMM (f3d_gemm_set_stream): resident-weight,
streaming or tiled kernel for the per-node layers (N = 128, 256).

The table lives in fiber3d_amd/tuning.json; the kernels module calls apply()
at import when it exists.  validate() checks a table's shape and ranges.
"""
import json
import os
import time

import numpy as np

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")
KERNELS = {"oseen": L.OSEEN, "rotlet": L.ROTLET, "stresslet": L.STRESSLET}
PRECS = {"f64": L.F64, "f32c": L.F32C}
LG_MIN, LG_MAX = 8, 22          # size buckets of the library (log2 sizes)
KS = (1, 2, 4)
WAVES = (2048, 4096, 8192, 16384, 32768)
GRID_NT = (4096, 16384, 65536)
GRID_NS = (1024, 4096, 16384, 65536)
LU_WAVES = {"f64": (1, 2, 4), "f32c": (2, 4, 8)}


def set_launch(kernel, precision, k, waves):
    rc = L.load().f3d_pairsum_set_launch(KERNELS[kernel], PRECS[precision], int(k), int(waves))
    L.check(rc, "f3d_pairsum_set_launch")


def set_launch_bucket(kernel, precision, lg_nt, lg_ns, k, waves):
    rc = L.load().f3d_pairsum_set_launch_bucket(KERNELS[kernel], PRECS[precision], int(lg_nt),
                                                int(lg_ns), int(k), int(waves))
    L.check(rc, "f3d_pairsum_set_launch_bucket")


def plan(kernel, precision, ns, nt):
    """(K, chunks, sources per chunk) f3d_pairsum would use."""
    import ctypes
    out = (ctypes.c_int32 * 3)()
    rc = L.load().f3d_pairsum_plan(KERNELS[kernel], PRECS[precision], int(ns), int(nt),
                                   ctypes.addressof(out))
    L.check(rc, "f3d_pairsum_plan")
    return tuple(out)


def lg(n):
    """ceil(log2 n) for n >= 1 (the library's bucket index before clamping)."""
    return int(max(0, int(n) - 1).bit_length())


def _nearest(points, a, b):
    return min(points, key=lambda p: (abs(p[0] - a) + abs(p[1] - b), p))


def expand_buckets(tuned):
    """{(lg_nt, lg_ns): entry} over the library's whole bucket range, each
    from the nearest tuned size."""
    pts = sorted(tuned)
    return {(a, b): tuned[_nearest(pts, a, b)]
            for a in range(LG_MIN, LG_MAX + 1) for b in range(LG_MIN, LG_MAX + 1)}


def validate(table):
    """Raise ValueError unless `table` is a well-formed tuning table whose
    every shape the library accepts."""
    if not isinstance(table, dict):
        raise ValueError("table must be a JSON object")
    for key, v in table.get("pairsum", {}).items():
        kernel, precision = key.split("/")
        if kernel not in KERNELS or precision not in PRECS:
            raise ValueError(f"unknown pairsum entry {key}")
        if v["k"] not in KS or not (64 <= v["waves"] <= 1 << 20):
            raise ValueError(f"bad pairsum shape {key}: {v}")
    for key, buckets in table.get("pairsum_buckets", {}).items():
        kernel, precision = key.split("/")
        if kernel not in KERNELS or precision not in PRECS:
            raise ValueError(f"unknown pairsum bucket entry {key}")
        for size, v in buckets.items():
            a, b = (int(s) for s in size.split("/"))
            if not (LG_MIN <= a <= LG_MAX and LG_MIN <= b <= LG_MAX):
                raise ValueError(f"bucket {key} {size} outside [{LG_MIN}, {LG_MAX}]")
            if v["k"] not in KS or not (64 <= v["waves"] <= 1 << 20):
                raise ValueError(f"bad bucket shape {key} {size}: {v}")
    for prec, v in table.get("lu", {}).items():
        if prec not in PRECS:
            raise ValueError(f"unknown LU precision {prec}")
        cap = 4 if prec == "f64" else 8
        if not (1 <= v["factor_waves_per_cu"] <= cap) or v["solve_staged"] not in (0, 1):
            raise ValueError(f"bad LU launch {prec}: {v}")
    for n, v in table.get("gemm", {}).items():
        if n not in ("128", "256") or v["stream"] not in (0, 1, 2):
            raise ValueError(f"bad GEMM entry {n}: {v}")
    return table


def apply(path=CACHE, table=None):
    """Install a tuning table (the saved one by default); returns it (empty
    if there is none)."""
    if table is None:
        if not os.path.exists(path):
            return {}
        with open(path) as fh:
            table = json.load(fh)
    validate(table)
    lib = L.load()
    for key, v in table.get("pairsum", {}).items():
        kernel, precision = key.split("/")
        set_launch(kernel, precision, v["k"], v["waves"])
    for key, buckets in table.get("pairsum_buckets", {}).items():
        kernel, precision = key.split("/")
        tuned = {tuple(int(s) for s in size.split("/")): v for size, v in buckets.items()}
        for (a, b), v in expand_buckets(tuned).items():
            set_launch_bucket(kernel, precision, a, b, v["k"], v["waves"])
    for prec, v in table.get("lu", {}).items():
        L.check(lib.f3d_lu_set_launch(PRECS[prec], int(v["factor_waves_per_cu"]),
                                      int(v["solve_staged"])), "f3d_lu_set_launch")
    for n, v in table.get("gemm", {}).items():
        L.check(lib.f3d_gemm_set_stream(int(n), int(v["stream"])), "f3d_gemm_set_stream")
    return table


def _time(fn, min_s=0.02, max_reps=200):
    """Seconds per call: one untimed call, then enough calls for ~min_s."""
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    one = max(time.perf_counter() - t0, 1e-6)
    reps = int(min(max_reps, max(3, min_s / one)))
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_pairsum(kernels=("oseen", "stresslet", "rotlet"), precisions=("f32c", "f64"),
                 grid_nt=GRID_NT, grid_ns=GRID_NS, ks=KS, waves=WAVES, seed=0, log=None):
    """Fastest (K, waves) per (kernel, precision) and tuned size; the
    65536 x 65536 entry is also the (kernel, precision) default."""
    import torch
    from .kernels import pairsum

    rng = np.random.default_rng(seed)
    nmax = max(max(grid_nt), max(grid_ns))
    X = torch.from_numpy(rng.uniform(0, 1, (nmax, 3))).cuda()
    F = torch.from_numpy(rng.normal(size=(nmax, 3))).cuda()
    NR = torch.nn.functional.normalize(torch.from_numpy(rng.normal(size=(nmax, 3))), dim=1).cuda()
    out = torch.empty((nmax, 3), dtype=torch.float64, device="cuda")
    defaults, buckets = {}, {}
    for kernel in kernels:
        for precision in precisions:
            key = f"{kernel}/{precision}"
            buckets[key] = {}
            for nt in grid_nt:
                for ns in grid_ns:
                    best = None
                    for k in ks:
                        for w in waves:
                            set_launch_bucket(kernel, precision, lg(nt), lg(ns), k, w)
                            kw = dict(src_nrm=NR[:ns] if kernel == "stresslet" else None,
                                      precision=precision, out=out[:nt])
                            t = _time(lambda: pairsum(kernel, X[:ns], F[:ns], X[:nt], **kw))
                            rate = nt * ns / t / 1e9
                            if best is None or rate > best[0]:
                                best = (rate, k, w)
                    set_launch_bucket(kernel, precision, lg(nt), lg(ns), best[1], best[2])
                    ent = {"k": best[1], "waves": best[2], "gpairs_s": round(best[0], 1)}
                    buckets[key][f"{lg(nt)}/{lg(ns)}"] = ent
                    if log:
                        log(f"{key} nt={nt} ns={ns}: {ent}")
                    if nt == ns == 65536:
                        defaults[key] = dict(ent)
    return defaults, buckets


def tune_pairsum_fibers(precisions=("f32c", "f64"), ks=KS, waves=WAVES, log=None):
    """The fiber -> fiber Oseen sum of the large workload (4096 fibers x 64
    nodes in a ball of radius 14, each fiber's own nodes excluded): its
    spatially compact tiles and exclusion ranges (the given order is kept)
    make the best shape differ from the unit cube's; tuned at 262144 x
    262144, bucket 18/18."""
    import torch
    from . import configs
    from .kernels import pairsum

    cfg = configs.config(4)
    fibers = configs.make_fibers(cfg.fibers, cfg.seed)
    x = np.concatenate(fibers)
    lens = np.array([len(f) for f in fibers])
    lo = np.repeat(np.cumsum(lens) - lens, lens).astype(np.int32)
    hi = (lo + np.repeat(lens, lens)).astype(np.int32)
    X = torch.from_numpy(x).cuda()
    F = torch.from_numpy(np.random.default_rng(1).normal(size=x.shape)).cuda()
    excl = (torch.from_numpy(lo).cuda(), torch.from_numpy(hi).cuda())
    out = torch.empty_like(X)
    n = x.shape[0]
    res = {}
    for precision in precisions:
        best = None
        for k in ks:
            for w in waves:
                set_launch_bucket("oseen", precision, lg(n), lg(n), k, w)
                t = _time(lambda: pairsum("oseen", X, F, X, excl=excl, precision=precision,
                                          out=out), min_s=0.05, max_reps=5)
                rate = n * n / t / 1e9
                if best is None or rate > best[0]:
                    best = (rate, k, w)
        set_launch_bucket("oseen", precision, lg(n), lg(n), best[1], best[2])
        ent = {"k": best[1], "waves": best[2], "gpairs_s": round(best[0], 1)}
        res[f"oseen/{precision}"] = {f"{lg(n)}/{lg(n)}": ent}
        if log:
            log(f"oseen/{precision} fibers n={n}: {ent}")
    return res


def _lu_workload(seed=0):
    """Fiber blocks of the stepper's shapes: 1000 fibers of 32 nodes
    (n = 128) and 512 of 64 nodes (n = 256), A_sq from the discretisation."""
    from . import configs
    from .system import FiberSystem
    out = []
    for nfib, N in ((1000, 32), (512, 64)):
        spec = configs.FiberSpec(nfib, N, 1.0, 6.0)
        S = FiberSystem(configs.make_fibers(spec, seed), None, lu_precision="f64")
        out.append(S)
    return out


def tune_lu(log=None):
    import torch
    from .dense import BatchedLU

    systems = _lu_workload()
    res = {}
    for prec in ("f64", "f32c"):
        best = None
        for w in LU_WAVES[prec]:
            L.check(L.load().f3d_lu_set_launch(PRECS[prec], w, 1), "f3d_lu_set_launch")
            lus = [BatchedLU(S.bs, S.A_buf, precision=prec, refine=0) for S in systems]
            t = _time(lambda: [lu.factor() for lu in lus])
            if best is None or t < best[0]:
                best = (t, w)
        timings = {}
        for staged in (1, 0):
            L.check(L.load().f3d_lu_set_launch(PRECS[prec], best[1], staged),
                    "f3d_lu_set_launch")
            lus = [BatchedLU(S.bs, S.A_buf, precision=prec, refine=0) for S in systems]
            xs = [torch.randn(S.bs.nrows, dtype=torch.float64, device="cuda") for S in systems]
            ys = [torch.empty_like(x) for x in xs]
            timings[staged] = _time(lambda: [lu.solve(x, out=y)
                                             for lu, x, y in zip(lus, xs, ys)])
        staged = 1 if timings[1] <= timings[0] else 0
        L.check(L.load().f3d_lu_set_launch(PRECS[prec], best[1], staged), "f3d_lu_set_launch")
        res[prec] = {"factor_waves_per_cu": best[1], "factor_ms": round(best[0] * 1e3, 4),
                     "solve_staged": staged, "solve_ms_staged": round(timings[1] * 1e3, 4),
                     "solve_ms_streaming": round(timings[0] * 1e3, 4)}
        if log:
            log(f"lu {prec}: {res[prec]}")
    return res


def tune_gemm(M=262144, K=256, log=None):
    import torch
    from . import learned

    res = {}
    g = torch.Generator().manual_seed(0)
    for N in (128, 256):
        A = torch.randn(M, K, generator=g).cuda()
        B = (torch.randn(K, N, generator=g) / 16).cuda()
        bias = torch.randn(N, generator=g).cuda()
        C = torch.empty(M, N, device="cuda")
        t = {}
        for mode in (2, 1, 0):
            L.check(L.load().f3d_gemm_set_stream(N, mode), "f3d_gemm_set_stream")
            t[mode] = _time(lambda: learned.gemm(A, B, C, epi=learned.EPI_BIAS_RELU, bias=bias))
        best = min(t, key=t.get)
        L.check(L.load().f3d_gemm_set_stream(N, best), "f3d_gemm_set_stream")
        res[str(N)] = {"stream": best, "ms_resident": round(t[2] * 1e3, 4),
                       "ms_stream": round(t[1] * 1e3, 4), "ms_tiled": round(t[0] * 1e3, 4),
                       "shape": [M, K, N]}
        if log:
            log(f"gemm N={N}: {res[str(N)]}")
    return res


PARTS = ("pairsum", "fibers", "lu", "gemm")


def tune_all(path=CACHE, log=print, parts=PARTS, base=None):
    """Tune `parts` and save the table; parts not re-tuned are kept from
    `base` (a table, e.g. the saved one)."""
    import torch
    table = dict(base or {})
    if "pairsum" in parts:
        defaults, buckets = tune_pairsum(log=log)
        table["pairsum"] = defaults
        old_fib = {k: {s: v for s, v in b.items() if s == "18/18"}
                   for k, b in table.get("pairsum_buckets", {}).items()}
        table["pairsum_buckets"] = buckets
        for k, b in old_fib.items():
            table["pairsum_buckets"].setdefault(k, {}).update(b)
    if "fibers" in parts:
        for k, b in tune_pairsum_fibers(log=log).items():
            table.setdefault("pairsum_buckets", {}).setdefault(k, {}).update(b)
    if "lu" in parts:
        table["lu"] = tune_lu(log=log)
    if "gemm" in parts:
        table["gemm"] = tune_gemm(log=log)
    table["device"] = torch.cuda.get_device_name(0)
    table["problem"] = ("pair sums: unit-cube points, n_targets x n_sources over "
                        f"{list(GRID_NT)} x {list(GRID_NS)}, and the 262144-node fiber -> "
                        "fiber Oseen sum with exclusion ranges (bucket 18/18); LU: 1000 "
                        "blocks of order 128 and 512 of order 256; GEMM: 262144 x 256 x N")
    validate(table)
    if path:
        with open(path, "w") as fh:
            json.dump(table, fh, indent=1)
    return table


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune launch shapes; writes a tuning table")
    ap.add_argument("path", nargs="?", default=CACHE)
    ap.add_argument("--only", nargs="+", choices=PARTS, default=list(PARTS),
                    help="re-tune these parts, keeping the others from the saved table")
    a = ap.parse_args()
    base = None
    if os.path.exists(CACHE):
        with open(CACHE) as fh:
            base = json.load(fh)
    tune_all(path=a.path, parts=tuple(a.only), base=base)
