"""The allocation-free pair sum (f3d_pairsum_noalloc) against f3d_pairsum,
both called through the C ABI: two calls in a row (the second smaller, so it
reuses scratch the first left behind), FP64 and FP32-compensated, with
exclusion ranges and the self-pair flag, plus a size whose default launch
shape needs more than the static scratch (fewer source chunks) and one that
cannot fit at all (rejected on the host before any launch)."""

import numpy as np
import pytest
import torch

from fiber3d_amd import _lib as L
from fiber3d_amd import kernels as FK

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32c": 1e-6}
PREC = {"f64": L.F64, "f32c": L.F32C}
KIND = {"oseen": L.OSEEN, "rotlet": L.ROTLET, "stresslet": L.STRESSLET}


def _case(rng, ns, nt):
    s = torch.from_numpy(rng.uniform(-1, 1, (ns, 3))).cuda()
    t = torch.from_numpy(rng.uniform(-1, 1, (nt, 3))).cuda()
    f = torch.from_numpy(rng.normal(size=(ns, 3))).cuda()
    n = torch.nn.functional.normalize(torch.from_numpy(rng.normal(size=(ns, 3))), dim=1).cuda()
    w = torch.from_numpy(rng.uniform(0.5, 1.5, ns)).cuda()
    return s, t, f, n, w


def _call(noalloc, kind, prec, flags, s, f, n, w, t, lo=None, hi=None):
    lib = L.load()
    nt, ns = t.shape[0], s.shape[0]
    out = torch.full((nt, 3), float("nan"), dtype=torch.float64, device="cuda")
    st = FK.ctypes_stream(torch.cuda.current_stream())
    args = [KIND[kind], PREC[prec], flags, ns, L.ptr(s), L.ptr(n if kind == "stresslet" else None),
            L.ptr(f), L.ptr(w), 3, 1, nt, L.ptr(t), 3, 1, L.ptr(lo), L.ptr(hi), 1.0, 5e-3, 1e-5,
            L.ptr(out), 3, 1]
    if noalloc:
        rc = lib.f3d_pairsum_noalloc(*args, st)
        L.check(rc, "f3d_pairsum_noalloc")
    else:
        nbytes = lib.f3d_pairsum_workspace_size(KIND[kind], PREC[prec], ns, nt)
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        rc = lib.f3d_pairsum(*args, L.ptr(ws), nbytes, st)
        L.check(rc, "f3d_pairsum")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", ["f64", "f32c"])
@pytest.mark.parametrize("kind", ["oseen", "stresslet", "rotlet"])
def test_noalloc_matches_pairsum_twice(kind, prec):
    rng = np.random.default_rng(11)
    for ns, nt in ((3000, 2500), (700, 450)):  # second call smaller
        s, t, f, n, w = _case(rng, ns, nt)
        ref = _call(False, kind, prec, 0, s, f, n, w, t)
        got = _call(True, kind, prec, 0, s, f, n, w, t)
        assert np.isfinite(got).all()
        assert _rel(got, ref) <= TOL[prec], (kind, prec, ns, nt, _rel(got, ref))


@pytest.mark.parametrize("prec", ["f64", "f32c"])
def test_noalloc_exclusion_and_self_flag(prec):
    rng = np.random.default_rng(12)
    for n_pts in (2048, 960):
        s, _, f, n, w = _case(rng, n_pts, 1)
        # same-fiber exclusion: fibers of 32 consecutive points
        idx = torch.arange(n_pts, dtype=torch.int32, device="cuda")
        lo = (idx // 32) * 32
        hi = torch.clamp(lo + 32, max=n_pts).to(torch.int32)
        ref = _call(False, "oseen", prec, 0, s, f, n, w, s, lo, hi)
        got = _call(True, "oseen", prec, 0, s, f, n, w, s, lo, hi)
        assert _rel(got, ref) <= TOL[prec]
        ref = _call(False, "stresslet", prec, L.EXCLUDE_SELF, s, f, n, None, s)
        got = _call(True, "stresslet", prec, L.EXCLUDE_SELF, s, f, n, None, s)
        assert _rel(got, ref) <= TOL[prec]


def test_noalloc_fewer_chunks_when_the_default_plan_does_not_fit():
    # a launch shape aiming at 2^18 waves splits 65536 sources over 262144
    # targets into 128 chunks, combined 4 per workgroup: 32 x 3 x 262144 x 8
    # bytes of partial sums, more than the static scratch, which then holds
    # fewer, longer chunks
    from fiber3d_amd import autotune
    lib = L.load()
    ns, nt = 65536, 262144
    rng = np.random.default_rng(13)
    s, t, f, n, w = _case(rng, ns, nt)
    try:
        for prec in ("f64", "f32c"):
            autotune.set_launch_bucket("oseen", prec, autotune.lg(nt), autotune.lg(ns), 2,
                                       1 << 18)
            assert autotune.plan("oseen", prec, ns, nt)[1] == 128
            assert lib.f3d_pairsum_workspace_size(L.OSEEN, PREC[prec], ns, nt) > (64 << 20)
            ref = _call(False, "oseen", prec, 0, s, f, n, None, t)
            got = _call(True, "oseen", prec, 0, s, f, n, None, t)
            assert _rel(got, ref) <= TOL[prec], (prec, _rel(got, ref))
    finally:
        autotune.apply()


def test_noalloc_rejects_sizes_beyond_the_scratch():
    # 2^21 stresslet sources in FP64 need 192 MiB of records alone; the host
    # check refuses the call before anything is launched
    lib = L.load()
    dummy = torch.zeros((1, 3), dtype=torch.float64, device="cuda")
    out = torch.zeros((1, 3), dtype=torch.float64, device="cuda")
    st = FK.ctypes_stream(torch.cuda.current_stream())
    rc = lib.f3d_pairsum_noalloc(L.STRESSLET, L.F64, 0, 1 << 21, L.ptr(dummy), L.ptr(dummy),
                                 L.ptr(dummy), None, 3, 1, 1, L.ptr(out), 3, 1, None, None, 1.0,
                                 5e-3, 1e-5, L.ptr(out), 3, 1, st)
    assert rc == L.EWORKSPACE
    assert b"static scratch" in lib.f3d_last_error()
