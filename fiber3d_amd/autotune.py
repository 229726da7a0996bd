"""Autotuner of the pair-sum launch shape.

For each (kernel, precision) it times f3d_pairsum over a grid of launch
shapes -- targets per lane K in {1, 2, 4} and the wavefront count the source
split aims for -- on a representative problem, keeps the fastest, installs it
with f3d_pairsum_set_launch and records it in a JSON cache
(fiber3d_amd/tuning.json by default).  `apply()` installs a saved table; the
kernels module calls it at import when the cache exists.
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


def set_launch(kernel, precision, k, waves):
    rc = L.load().f3d_pairsum_set_launch(KERNELS[kernel], PRECS[precision], int(k), int(waves))
    L.check(rc, "f3d_pairsum_set_launch")


def apply(path=CACHE):
    """Install a saved tuning table; returns it (empty if there is none)."""
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        table = json.load(fh)
    for key, v in table.get("pairsum", {}).items():
        kernel, precision = key.split("/")
        set_launch(kernel, precision, v["k"], v["waves"])
    return table


def tune_pairsum(n=65536, kernels=("oseen", "stresslet", "rotlet"), precisions=("f32c", "f64"),
                 ks=(1, 2, 4), waves=(4096, 8192, 16384), reps=3, path=CACHE, seed=0):
    import torch
    from .kernels import pairsum

    rng = np.random.default_rng(seed)
    X = torch.from_numpy(rng.uniform(0, 1, (n, 3))).cuda()
    F = torch.from_numpy(rng.normal(size=(n, 3))).cuda()
    NR = torch.nn.functional.normalize(torch.from_numpy(rng.normal(size=(n, 3))), dim=1).cuda()
    out = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    table = {"pairsum": {}, "problem": f"{n} x {n} unit cube", "device":
             torch.cuda.get_device_name(0)}
    for kernel in kernels:
        for precision in precisions:
            best = None
            for k in ks:
                for w in waves:
                    set_launch(kernel, precision, k, w)
                    kw = dict(src_nrm=NR if kernel == "stresslet" else None, precision=precision,
                              out=out)
                    pairsum(kernel, X, F, X, **kw)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        pairsum(kernel, X, F, X, **kw)
                    torch.cuda.synchronize()
                    rate = n * n * reps / (time.perf_counter() - t0) / 1e9
                    if best is None or rate > best[0]:
                        best = (rate, k, w)
            set_launch(kernel, precision, best[1], best[2])
            table["pairsum"][f"{kernel}/{precision}"] = {"k": best[1], "waves": best[2],
                                                         "gpairs_s": round(best[0], 1)}
    if path:
        with open(path, "w") as fh:
            json.dump(table, fh, indent=1)
    return table


if __name__ == "__main__":
    import sys
    print(json.dumps(tune_pairsum(path=sys.argv[1] if len(sys.argv) > 1 else CACHE), indent=1))
