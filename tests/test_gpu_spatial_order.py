"""GPU parity of the pair sums run in spatial order.

Oseen and rotlet calls with at least 65536 sources and targets visit both in
Morton order (a device radix sort); FP32C evaluates far tiles on
tile-relative FP32 coordinates, FP64 far tiles skip the near-pair vote.
F3D_PAIRSUM_SORT=1 forces the sort at the small sizes here, so each case runs
both ways and meets the FP64 oracle at the bars (relative 2-norm: 1e-6
FP32C, 1e-10 FP64): ragged sizes, distinct sources and targets, the
self-sum flag (one shared order, slot i excluding source i), and clustered
points where most tiles are far.  Calls with exclusion ranges keep the given
order, and the stresslet never sorts (its FP32 terms cancel, and spatially
grouped tile sums raise its error by about 15%): their cases check the far
tiles of the given order.
"""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import kernels as FK
from oracle import kernels as K

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32c": 1e-6}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(params=["0", "1"], ids=["given_order", "spatial_order"])
def order(request, monkeypatch):
    monkeypatch.setenv("F3D_PAIRSUM_SORT", request.param)
    return request.param


def _clusters(rng, n, centres, size):
    c = centres[rng.integers(0, len(centres), size=n)]
    return c + rng.normal(scale=size, size=(n, 3))


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("kernel", ["oseen", "rotlet", "stresslet"])
@pytest.mark.parametrize("ns,nt", [(1500, 1300), (333, 2049), (65, 1), (1, 7)])
def test_sorted_sums_match_oracle(order, kernel, ns, nt, precision):
    print(f"case spatial {order} {kernel} {precision} ns={ns} nt={nt}", flush=True)
    rng = np.random.default_rng(ns * 31 + nt)
    centres = rng.uniform(-20, 20, size=(12, 3))
    s = _clusters(rng, ns, centres, 0.7)
    t = _clusters(rng, nt, centres, 0.7)
    d = rng.normal(size=(ns, 3))
    n = rng.normal(size=(ns, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    if kernel == "oseen":
        ref = K.oseen_kernel_source_target(s, t, d)
    elif kernel == "rotlet":
        ref = K.rotlet_kernel_source_target(s, t, d)
    else:
        ref = K.stresslet_kernel_source_target(s, t, n, d)
    got = FK.pairsum(kernel, _dev(s), _dev(d), _dev(t), src_nrm=_dev(n), precision=precision)
    torch.cuda.synchronize()
    assert rel2(got.cpu().numpy(), ref) < TOL[precision]


def test_sorted_exclusion_weights_soa_accumulate(order):
    """Fibers (exclusion ranges over each fiber's own nodes): the given
    order is kept, and the far fibers' tiles skip the exclusion test."""
    print(f"case spatial fibers {order}", flush=True)
    rng = np.random.default_rng(4)
    nfib, N = 60, 24
    centres = rng.uniform(-10, 10, size=(nfib, 3))
    tang = rng.normal(size=(nfib, 3))
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    sgrid = np.linspace(-0.5, 0.5, N)
    s = (centres[:, None, :] + sgrid[None, :, None] * tang[:, None, :]).reshape(-1, 3)
    f = rng.normal(size=(nfib * N, 3))
    w = rng.uniform(0.5, 1.5, size=nfib * N)
    lo = np.repeat(np.arange(nfib) * N, N).astype(np.int32)
    hi = lo + N
    ref = K.oseen_kernel_source_target(s, s, f * w[:, None], exclude=(lo, hi))
    S = _dev(s.T.copy()).T
    F = _dev(f.T.copy()).T
    out = torch.full((3, nfib * N), 1.5, dtype=torch.float64, device="cuda").T
    FK.pairsum("oseen", S, F, S, src_w=_dev(w), excl=(torch.from_numpy(lo).cuda(),
               torch.from_numpy(hi).cuda()), precision="f32c", out=out, accumulate=True)
    torch.cuda.synchronize()
    assert rel2(out.cpu().numpy().ravel() - 1.5, ref) < TOL["f32c"]


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_sorted_self_sum(order, precision):
    print(f"case spatial self {order} {precision}", flush=True)
    rng = np.random.default_rng(23)
    x = _clusters(rng, 700, rng.uniform(-8, 8, size=(5, 3)), 1.0)
    f = rng.normal(size=(700, 3))
    i = np.arange(700, dtype=np.int32)
    ref = K.rotlet_kernel_source_target(x, x, f, reg=0.0, epsilon_distance=0.0,
                                        exclude=(i, i + 1))
    X = _dev(x)
    got = FK.pairsum("rotlet", X, _dev(f), X, reg=0.0, eps=0.0, precision=precision,
                     exclude_self=True).cpu().numpy().ravel()
    assert rel2(got, ref) < TOL[precision]


def test_far_clusters_and_default_policy():
    """Two clusters far apart (every tile of one is far from every wave of
    the other), and a 65536-point self-sum, the size from which the sort is
    on by default, on 200 sampled rows: both meet the oracle."""
    print("case spatial far clusters", flush=True)
    rng = np.random.default_rng(11)
    a = rng.normal(scale=0.3, size=(3000, 3))
    b = rng.normal(scale=0.3, size=(2500, 3)) + np.array([40.0, 0.0, 0.0])
    f = rng.normal(size=(3000, 3))
    ref = K.oseen_kernel_source_target(a, b, f)
    got = FK.pairsum("oseen", _dev(a), _dev(f), _dev(b), precision="f32c").cpu().numpy()
    assert rel2(got, ref) < TOL["f32c"]
    n = 65536
    x = rng.uniform(0, 1, size=(n, 3))
    g = rng.normal(size=(n, 3))
    X, G = _dev(x), _dev(g)
    rows = rng.choice(n, size=200, replace=False)
    ref = K.oseen_kernel_source_target(x, x[rows], g, exclude=(rows.astype(np.int32),
                                                                 rows.astype(np.int32) + 1))
    for precision in ("f32c", "f64"):
        sorted_ = FK.pairsum("oseen", X, G, X, precision=precision,
                             exclude_self=True).cpu().numpy()
        assert rel2(sorted_[rows], ref) < TOL[precision]
