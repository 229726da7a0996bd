"""GPU batched LU / solve / block GEMV against LAPACK (SciPy) on the host."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from conftest import rel2
from fiber3d_amd.dense import BatchedLU, BlockGemv, BlockSet, DenseGemv

pytestmark = pytest.mark.gpu


def _mats(sizes, seed, cond_shift=4.0):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=(n, n)) + cond_shift * np.eye(n) for n in sizes]


SIZES = [64, 96, 128, 192, 256, 37, 1, 100, 255, 17]


def test_lu_factor_f64_matches_lapack():
    print("case lu f64 factor", flush=True)
    mats = _mats(SIZES, 0)
    bs = BlockSet(SIZES)
    A = bs.pack(mats)
    lu = BatchedLU(bs, A, precision="f64")
    torch.cuda.synchronize()
    LU = bs.unpack(lu.lu64)
    piv = lu.piv.cpu().numpy()
    assert lu.info.cpu().numpy().tolist() == [0] * len(SIZES)
    for b, M in enumerate(mats):
        lu_ref, piv_ref = sla.lu_factor(M)
        o = bs.x_off_h[b]
        np.testing.assert_array_equal(piv[o:o + SIZES[b]], piv_ref)
        assert rel2(LU[b], lu_ref) < 1e-12


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_lu_solve_batched(precision):
    print(f"case lu solve {precision}", flush=True)
    mats = _mats(SIZES, 1)
    bs = BlockSet(SIZES)
    A = bs.pack(mats)
    lu = BatchedLU(bs, A, precision=precision, refine=10, tol=1e-14)
    rng = np.random.default_rng(2)
    xs = [rng.normal(size=n) for n in SIZES]
    b = np.concatenate([M @ x for M, x in zip(mats, xs)])
    y = lu.solve(torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    tol = 1e-12
    for k, x in enumerate(xs):
        o = bs.x_off_h[k]
        assert rel2(yh[o:o + SIZES[k]], x) < tol, (k, SIZES[k])
    if precision == "f32c":
        it = lu.iters.cpu().numpy()
        assert it.max() >= 1 and it.max() < 10
    # in-place (y aliases x)
    bb = torch.from_numpy(b).cuda()
    lu.solve(bb, out=bb)
    torch.cuda.synchronize()
    assert rel2(bb.cpu().numpy(), np.concatenate(xs)) < tol


def test_lu_zero_pivot_info():
    M = np.eye(8)
    M[3, 3] = 0.0
    M[3, :] = 0.0
    bs = BlockSet([8, 8])
    lu = BatchedLU(bs, bs.pack([np.eye(8) * 2.0, M]), precision="f64")
    torch.cuda.synchronize()
    assert lu.info.cpu().numpy().tolist() == [0, 4]


def test_block_gemv():
    print("case block gemv", flush=True)
    rng = np.random.default_rng(4)
    ms = [12, 256, 64, 130]
    ks = [16, 200, 256, 3]
    mats = [rng.normal(size=(m, k)) for m, k in zip(ms, ks)]
    x_off = np.concatenate([[0], np.cumsum(ks)[:-1]])
    y_off = np.concatenate([[0], np.cumsum(ms)[:-1]])
    g = BlockGemv(ms, ks, x_off, y_off)
    M = g.pack(mats)
    x = rng.normal(size=sum(ks))
    y0 = rng.normal(size=sum(ms))
    y = torch.from_numpy(y0.copy()).cuda()
    g(M, torch.from_numpy(x).cuda(), y, alpha=2.0, beta=-0.5)
    torch.cuda.synchronize()
    ref = np.concatenate([2.0 * Mb @ x[xo:xo + k] for Mb, xo, k in zip(mats, x_off, ks)]) - 0.5 * y0
    assert rel2(y.cpu().numpy(), ref) < 1e-13


def test_batched_lu_on_real_fiber_blocks_config2_shape():
    """Workload 2 shape: fibers with N in {16, 24, 32, 48, 64} (n = 4N up to
    256), A_sq assembled from the fiber discretisation; FP32 LU + FP64
    refinement to 1e-12 against the FP64 LU in one batch."""
    print("case batched lu config2-like", flush=True)
    from fiber3d_amd import configs
    from fiber3d_amd.system import FiberSystem
    spec = configs.FiberSpec(160, [16, 24, 32, 48, 64], 1.0, 8.0)
    fibers = configs.make_fibers(spec, seed=2)
    S = FiberSystem(fibers, None, lu_precision="f32c", lu_refine=10, lu_tol=1e-14)
    rng = np.random.default_rng(5)
    xs = torch.from_numpy(rng.normal(size=S.bs.nrows)).cuda()
    b = torch.empty_like(xs)
    S.gA(S.A_buf, xs, b)          # b = A_sq x for every block
    y32 = S.lu.solve(b)
    lu64 = BatchedLU(S.bs, S.A_buf, precision="f64")
    y64 = lu64.solve(b)
    torch.cuda.synchronize()
    assert S.lu.info.cpu().numpy().max() == 0
    e32 = rel2(y32.cpu().numpy(), xs.cpu().numpy())
    e64 = rel2(y64.cpu().numpy(), xs.cpu().numpy())
    print(f"forward error f32+IR {e32:.2e}  f64 {e64:.2e}  max sweeps {int(S.lu.iters.max())}")
    # FP64 refinement drives every block's relative residual to 1e-12 or below
    Ay = torch.empty_like(b)
    S.gA(S.A_buf, y32, Ay)
    r = (b - Ay).cpu().numpy()
    bh = b.cpu().numpy()
    for k in range(S.bs.nb):
        o, n = S.bs.x_off_h[k], S.bs.sizes[k]
        assert np.linalg.norm(r[o:o + n]) <= 1e-12 * np.linalg.norm(bh[o:o + n]), k
    # forward error stays at the conditioning level of the FP64 factorisation
    assert e32 < max(1e4 * e64, 1e-9)


@pytest.mark.parametrize("kernel", ["mfma", "vector"])
def test_lu_factor_f32_is_a_factorisation(kernel, monkeypatch):
    """The equilibrated FP32 factors satisfy P (Dr A Dc) = L U to FP32
    accuracy, with ipiv and perm consistent (blocked MFMA kernel and the
    register-resident VALU kernel; orders not divisible by 32 included)."""
    print(f"case lu f32 factor {kernel}", flush=True)
    if kernel == "vector":
        monkeypatch.setenv("F3D_LU32_KERNEL", "vector")
    else:
        monkeypatch.delenv("F3D_LU32_KERNEL", raising=False)
    mats = _mats(SIZES, 7)
    bs = BlockSet(SIZES)
    lu = BatchedLU(bs, bs.pack(mats), precision="f32c")
    torch.cuda.synchronize()
    assert lu.info.cpu().numpy().tolist() == [0] * len(SIZES)
    LU = bs.unpack(lu.lu32.double())
    piv, perm = lu.piv.cpu().numpy(), lu.perm.cpu().numpy()
    rs, cs = lu.rs.cpu().numpy(), lu.cs.cpu().numpy()
    for b, M in enumerate(mats):
        n, o = SIZES[b], bs.x_off_h[b]
        As = rs[o:o + n, None] * M * cs[None, o:o + n]
        p = perm[o:o + n]
        ref = np.arange(n)
        for j in range(n):
            q = piv[o + j]
            assert j <= q < n
            ref[[j, q]] = ref[[q, j]]
        np.testing.assert_array_equal(p, ref)
        L = np.tril(LU[b], -1) + np.eye(n)
        U = np.triu(LU[b])
        assert np.abs(L).max() <= 1.0 + 1e-6          # partial pivoting
        err = np.abs(As[p] - L @ U).max() / np.abs(As).max()
        assert err < 2e-5, (n, err)


@pytest.mark.parametrize("count,m,offset", [(1, 3072, 0), (3, 1806, 0), (2, 7, 0), (2, 130, 1),
                                            (1, 1, 0)])
def test_dense_gemv_batched(count, m, offset):
    """Row-per-wave dense GEMV against NumPy in FP64: even and odd sizes, and
    an x that is not 16-byte aligned (the one-double-per-lane path)."""
    print(f"case dense gemv count={count} m={m} offset={offset}", flush=True)
    rng = np.random.default_rng(m + count)
    M = rng.normal(size=(count, m, m))
    x = rng.normal(size=count * m)
    g = DenseGemv(torch.from_numpy(M).cuda())
    xbuf = torch.zeros(count * m + offset, dtype=torch.float64, device="cuda")
    xbuf[offset:] = torch.from_numpy(x)
    y = torch.full((count * m,), float("nan"), dtype=torch.float64, device="cuda")
    g(xbuf[offset:], y)
    torch.cuda.synchronize()
    ref = np.einsum("bij,bj->bi", M, x.reshape(count, m)).ravel()
    assert rel2(y.cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize("count,m,offset", [(1, 3072, 0), (3, 1806, 0), (2, 7, 0), (2, 132, 1),
                                            (1, 1, 0)])
def test_dense_gemv_batched_f32_entries(count, m, offset):
    """FP32 matrix entries, FP64 vector and arithmetic: equal to the FP64
    product of the FP32-rounded matrix to FP64 rounding, on the four-floats-
    per-lane path (m a multiple of 4, aligned), odd sizes and an x that is not
    16-byte aligned (one entry per lane)."""
    print(f"case dense gemv f32 count={count} m={m} offset={offset}", flush=True)
    rng = np.random.default_rng(m + count + 1)
    M32 = rng.normal(size=(count, m, m)).astype(np.float32)
    x = rng.normal(size=count * m)
    g = DenseGemv(torch.from_numpy(M32).cuda())
    xbuf = torch.zeros(count * m + offset, dtype=torch.float64, device="cuda")
    xbuf[offset:] = torch.from_numpy(x)
    y = torch.full((count * m,), float("nan"), dtype=torch.float64, device="cuda")
    g(xbuf[offset:], y)
    torch.cuda.synchronize()
    ref = np.einsum("bij,bj->bi", M32.astype(np.float64), x.reshape(count, m)).ravel()
    assert rel2(y.cpu().numpy(), ref) < 1e-13
