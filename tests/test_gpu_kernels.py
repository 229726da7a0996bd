"""GPU parity of the pair-sum and periphery kernels against the FP64 oracle.

Bars: FP64 path relative 2-norm <= 1e-10, FP32-compensated path <= 1e-6.
"""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import kernels as FK
from fiber3d_amd.geometry import shape_gallery_sphere
from oracle import kernels as K
from oracle import periphery as P

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32c": 1e-6}
ORACLE = {
    "oseen": lambda s, t, d, n, **kw: K.oseen_kernel_source_target(s, t, d, **kw),
    "rotlet": lambda s, t, d, n, **kw: K.rotlet_kernel_source_target(s, t, d, **kw),
    "stresslet": lambda s, t, d, n, **kw: K.stresslet_kernel_source_target(s, t, n, d, **kw),
}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _problem(seed, ns, nt, coincident=5, near=True):
    rng = np.random.default_rng(seed)
    s = rng.uniform(-1, 1, size=(ns, 3))
    t = rng.uniform(-1, 1, size=(nt, 3))
    k = min(coincident, ns, nt)
    t[:k] = s[:k]  # exact coincident pairs exercise the regularisation
    if near:
        # 0 < r < eps: regularised, d kept.  Only for FP64: with FP32
        # coordinates of size ~1 such a d = t - s carries ~1e-2 relative error
        # (a limit of the FP32 path, stated in its documentation).
        t[k:2 * k] = s[:k] + 3e-6
    d = rng.normal(size=(ns, 3))
    n = rng.normal(size=(ns, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return s, t, d, n


@pytest.mark.parametrize("kernel", ["oseen", "rotlet", "stresslet"])
@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("ns,nt", [(1500, 1300), (1, 7), (65, 1), (333, 2049)])
def test_pairsum_matches_oracle(kernel, precision, ns, nt):
    print(f"case pairsum {kernel} {precision} ns={ns} nt={nt}", flush=True)
    s, t, d, n = _problem(ns * 7 + nt, ns, nt, near=(precision == "f64"))
    eta = 0.8
    ref = ORACLE[kernel](s, t, d, n, eta=eta)
    got = FK.pairsum(kernel, _dev(s), _dev(d), _dev(t), src_nrm=_dev(n), eta=eta,
                     precision=precision)
    torch.cuda.synchronize()
    assert rel2(got.cpu().numpy(), ref) < TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_pairsum_soa_weights_exclusion_accumulate(precision):
    print(f"case pairsum soa/excl {precision}", flush=True)
    rng = np.random.default_rng(3)
    nfib, N = 40, 32
    s = rng.uniform(-2, 2, size=(nfib * N, 3))
    f = rng.normal(size=(nfib * N, 3))
    w = rng.uniform(0.5, 1.5, size=nfib * N)
    lo = np.repeat(np.arange(nfib) * N, N).astype(np.int32)
    hi = lo + N
    ref = K.oseen_kernel_source_target(s, s, f * w[:, None], exclude=(lo, hi))
    # SoA buffers (3, N) passed as their (N, 3) transposes; output SoA too
    S = _dev(s.T.copy()).T
    F = _dev(f.T.copy()).T
    out = torch.full((3, nfib * N), 1.5, dtype=torch.float64, device="cuda").T
    FK.pairsum("oseen", S, F, S, src_w=_dev(w), excl=(torch.from_numpy(lo).cuda(),
               torch.from_numpy(hi).cuda()), precision=precision, out=out, accumulate=True)
    torch.cuda.synchronize()
    assert rel2(out.cpu().numpy().ravel() - 1.5, ref) < TOL[precision]


def test_pairsum_empty_sources_and_targets():
    t = _dev(np.ones((5, 3)))
    out = FK.pairsum("oseen", _dev(np.zeros((0, 3))), _dev(np.zeros((0, 3))), t)
    assert torch.count_nonzero(out).item() == 0
    out2 = FK.pairsum("oseen", t, t, _dev(np.zeros((0, 3))))
    assert out2.shape == (0, 3)


def test_dropin_names_match_oracle():
    print("case dropin names", flush=True)
    s, t, d, n = _problem(5, 700, 500)
    np.testing.assert_allclose(FK.oseen_kernel_source_target_numba(s.ravel(), t.ravel(), d.ravel()),
                               K.oseen_kernel_source_target(s, t, d), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(FK.rotlet_kernel_source_target_numba(s, t, d, eta=2.0),
                               K.rotlet_kernel_source_target(s, t, d, eta=2.0), rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_allclose(FK.stresslet_kernel_source_target_numba(s, t, n, d),
                               K.stresslet_kernel_source_target(s, t, n, d), rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_allclose(FK.stresslet_kernel_times_normal_times_density_numba(s, n, d),
                               K.stresslet_kernel_times_normal_times_density(s, n, d), rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_allclose(FK.complementary_kernel_times_density_numba(s, n, d),
                               K.complementary_kernel_times_density(s, n, d), rtol=1e-12)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_periphery_apply_matches_dense_operator(precision):
    print(f"case periphery {precision}", flush=True)
    N, R = 800, 5.0
    x, nout = shape_gallery_sphere(N, R)
    nrm = -nout
    w = np.full(N, 4 * np.pi * R * R / N)
    ex, ey, ez = P.singularity_subtraction_vectors(x, nrm, w)
    rng = np.random.default_rng(8)
    q = rng.normal(size=(N, 3))
    add = rng.normal(size=(N, 3))
    M = P.shell_matrix(x, nrm, w, ex, ey, ez)
    ref = M @ q.ravel() + add.ravel()
    got = FK.periphery_apply(_dev(x), _dev(nrm), _dev(w), _dev(ex), _dev(ey), _dev(ez), _dev(q),
                             add=_dev(add), precision=precision)
    torch.cuda.synchronize()
    assert rel2(got.cpu().numpy(), ref) < TOL[precision]


def test_large_oseen_f32c_against_f64_and_oracle_rows():
    # Full raw-sweep size: 65536 points in the unit cube, self term excluded.
    print("case large oseen 65536", flush=True)
    rng = np.random.default_rng(1)
    N = 65536
    x = rng.uniform(0, 1, size=(N, 3))
    f = rng.normal(size=(N, 3))
    X, F = _dev(x), _dev(f)
    idx = torch.arange(N, dtype=torch.int32, device="cuda")
    excl = (idx, idx + 1)
    u64 = FK.pairsum("oseen", X, F, X, excl=excl, precision="f64")
    u32 = FK.pairsum("oseen", X, F, X, excl=excl, precision="f32c")
    torch.cuda.synchronize()
    assert rel2(u32.cpu().numpy(), u64.cpu().numpy()) < 1e-6
    # oracle on a sample of 96 targets against all sources
    rows = rng.choice(N, size=96, replace=False)
    ref = K.oseen_kernel_source_target(x, x[rows], f, exclude=(rows, rows + 1))
    assert rel2(u64.cpu().numpy()[rows], ref.reshape(-1, 3)) < 1e-10
    # linearity
    u2 = FK.pairsum("oseen", X, 2.0 * F - F, X, excl=excl, precision="f64")
    torch.cuda.synchronize()
    assert rel2(u2.cpu().numpy(), u64.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("kernel", ["oseen", "rotlet", "stresslet"])
@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_padding_sources_are_masked_without_regularisation(kernel, precision):
    """With eps = reg = 0 nothing is regularised.  100 sources leave 28
    zero-density padding records in the last 64-source tile, placed at the
    origin (FP64) or at the bounding box's minimum corner (FP32C, quantised
    0); a target sits exactly there.  The reference's sum over the real
    sources is finite at that target, and so must ours be."""
    print(f"case padding {kernel} {precision}", flush=True)
    rng = np.random.default_rng(21)
    s = rng.uniform(0.1, 1.0, size=(100, 3))
    t = np.vstack([np.zeros((1, 3)), rng.uniform(0.1, 1.0, size=(40, 3))])
    d = rng.normal(size=(100, 3))
    n = rng.normal(size=(100, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    ref = ORACLE[kernel](s, t, d, n, reg=0.0, epsilon_distance=0.0)
    assert np.isfinite(ref).all()
    got = FK.pairsum(kernel, _dev(s), _dev(d), _dev(t), src_nrm=_dev(n), reg=0.0, eps=0.0,
                     precision=precision).cpu().numpy().ravel()
    assert np.isfinite(got).all()
    assert rel2(got, ref) < TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_self_sum_drops_the_diagonal_without_regularisation(precision):
    """The stresslet self-sum (j != i) with eps = reg = 0: the i == j pair is
    dropped by the self-pair flag rather than relying on a regularised
    r = 0 term."""
    rng = np.random.default_rng(22)
    x = rng.uniform(-1, 1, size=(300, 3))
    n = rng.normal(size=(300, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    q = rng.normal(size=(300, 3))
    ref = K.stresslet_kernel_times_normal_times_density(x, n, q, reg=0.0, epsilon_distance=0.0)
    assert np.isfinite(ref).all()
    X = _dev(x)
    got = FK.pairsum("stresslet", X, _dev(q), X, src_nrm=_dev(n), reg=0.0, eps=0.0,
                     precision=precision, exclude_self=True).cpu().numpy().ravel()
    assert rel2(got, ref) < TOL[precision]
