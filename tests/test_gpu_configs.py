"""Parity at the full sizes of the stated workloads.

  * large: 4096 fibers x 64 nodes (length 1.5, ball radius 14), a 32768-node
    periphery of radius 16 and 8 rigid spheres of 600 nodes;
  * time series: 1000 fibers x 32 nodes, an 8000-node periphery of radius 10;
  * both: the GPU's A_sq blocks, right-hand side, FP64 matvec and fiber
    preconditioner against the sampled-rows oracle (oracle/sampled.py: C
    restatement of the reference's pair loops for the O(N^2) sums) on
    sampled fibers, sampled shell nodes and every body row; the
    FP32-compensated matvec against the FP64 one over the whole vector; one
    plain FP64 GMRES solve whose true residual, recomputed in FP64 on the GPU
    and checked on the oracle's rows, is <= 1e-8;
  * batched preconditioner: all 2048 fibers of N in {16, 24, 32, 48, 64};
  * raw sweep: the stresslet at N = M = 65536;
  * the small workload with Chebyshev-Lobatto nodes and spectral
    differentiation matrices.
Bars: 1e-6 relative 2-norm for FP32-compensated; for the FP64 path 1e-10, or
ten times the reference's own FP64 rounding floor where that is larger (the
assembled fiber blocks: derivatives of straight fibers far from the origin,
and spectral differentiation matrices, amplify rounding; the floor is measured
in each test by translating the fibers, which changes the exact blocks not at
all).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd import kernels as FK
from fiber3d_amd.dense import BatchedLU
from fiber3d_amd.gmres import gmres
from fiber3d_amd.system import FiberSystem, Shell
from oracle import build as obuild
from oracle import kernels as K
from oracle import sampled as OS
from oracle import tstep as OT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _clib():
    if not os.path.exists(obuild.OUT):
        obuild.build(verbose=False)


def _block(S, k):
    n = int(S.bs.sizes[k])
    o = int(S.bs.a_off_h[k])
    return S.A_buf[o:o + n * n].view(n, n).T.cpu().numpy()


def _workload(cfg_id, n_fib_rows, n_shell_rows, with_bodies):
    cfg = configs.config(cfg_id)
    spec = cfg.fibers
    print(f"case workload {cfg_id}: building {spec.num_fibers} fibers", flush=True)
    fibers = configs.make_fibers(spec, cfg.seed)
    xs, ns, ws = configs.make_periphery(cfg.periphery_nodes, cfg.periphery_radius)
    bodies = configs.make_bodies(8, 0.5, 600, spec.ball_radius, fibers=fibers, seed=cfg.seed) \
        if with_bodies else None
    lengths = np.full(len(fibers), spec.length)
    kw = dict(eta=cfg.eta, dt=cfg.dt, E=spec.E, epsilon=spec.epsilon, lengths=lengths,
              bodies=bodies)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws, eta=cfg.eta), **kw)
    rng = np.random.default_rng(cfg.seed)
    rows_f = sorted(rng.choice(len(fibers), n_fib_rows, replace=False).tolist())
    rows_s = sorted(rng.choice(cfg.periphery_nodes, n_shell_rows, replace=False).tolist())
    ora = OS.SampledHydro(fibers, (xs, ns, ws), fiber_rows=rows_f, shell_rows=rows_s, **kw)
    print(f"case workload {cfg_id}: system size {gpu.size}", flush=True)
    return gpu, ora


SHIFT = np.array([1e-3, -7e-4, 5e-4])


def _asq(ora, k, y):
    from oracle import fiber as OF
    fm = ora.fm[y.shape[0]]
    L = ora.lengths[k]
    A = OF.form_linear_operator(fm, y, L, ora.E, ora.dt, ora.c0, ora.c1, ora.penalty)
    R = OF.compute_RHS(fm, y, L, ora.dt, ora.c0, ora.c1, ora.penalty)
    return OF.apply_BC_rectangular(fm, y, L, ora.E, A, R)[0]


def _floors(ora, ks, x, y_rows):
    """Rounding floor of the reference's own FP64 assembly.  A fiber's A_sq
    depends on its nodes only through their derivatives, so translating the
    fiber (by 1e-3) changes it only by rounding -- chiefly of D_k x, whose
    noise grows with the coordinates' size and the norm of D_k (large for
    the spectral matrices).  Returns, per fiber k, the relative 2-norm of
    that change, and the size of the change it makes in the sampled matvec
    rows y_rows for the vector x, relative to ||y_rows||."""
    blk, d = {}, []
    for k in ks:
        xk = ora.fibers[k]
        A0, A1 = _asq(ora, k, xk), _asq(ora, k, xk + SHIFT)
        blk[k] = rel2(A1, A0)
        o = ora.fib0 + 4 * ora.offs[k]
        d.append((A1 - A0) @ x[o:o + A0.shape[0]])
    return blk, np.linalg.norm(np.concatenate(d)) / np.linalg.norm(y_rows)


def _bar(floor):
    """The FP64 bar: 1e-10, or ten times the reference's own rounding floor
    where that is larger."""
    return max(1e-10, 10.0 * floor)


def _check_workload(gpu, ora):
    assert gpu.size == ora.size
    idx = ora.row_index()
    idx_t = torch.from_numpy(idx).cuda()
    e_rhs = rel2(gpu.rhs[idx_t].cpu().numpy(), ora.rhs())
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.normal(size=gpu.size)).cuda()
    y64 = gpu.matvec(x, precision="f64")
    y32 = gpu.matvec(x, precision="f32c")
    torch.cuda.synchronize()
    xh = x.cpu().numpy()
    y_ora = ora.matvec(xh)
    fl_blk, fl_mv = _floors(ora, ora.fiber_rows, xh, y_ora)
    e_blk = max(rel2(_block(gpu, k), ora.A_sq[k]) / _bar(fl_blk[k]) for k in ora.fiber_rows)
    e_mv = rel2(y64[idx_t].cpu().numpy(), y_ora)
    e_32 = rel2(y32.cpu().numpy(), y64.cpu().numpy())
    print(f"A_sq error / bar {e_blk:.3f} (floors up to {max(fl_blk.values()):.2e}); "
          f"rhs {e_rhs:.3e}; matvec f64 {e_mv:.3e} (floor {fl_mv:.2e}); "
          f"f32c vs f64 {e_32:.3e}", flush=True)
    assert e_blk <= 1.0
    assert e_rhs < 1e-10
    assert e_mv < _bar(fl_mv)
    assert e_32 < 1e-6
    # fiber blocks of the preconditioner: LAPACK solves with the oracle's A_sq
    z = gpu.precond(x).cpu().numpy()
    for k in ora.fiber_rows:
        o = ora.fib0 + 4 * ora.offs[k]
        n = 4 * ora.sizes[k]
        assert rel2(z[o:o + n], sla.solve(ora.A_sq[k], xh[o:o + n])) < 1e-10, k
    # one plain FP64 solve to 1e-8 (the reference's zero initial guess)
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-8, restart=150, maxiter=200)
    torch.cuda.synchronize()
    print(f"plain solve: {res.iterations} iterations, residual {res.residual:.3e}", flush=True)
    assert res.info == 0 and res.residual <= 1e-8
    bnorm = float(torch.linalg.norm(gpu.rhs))
    r = gpu.rhs - gpu.matvec(res.x, precision="f64")
    assert float(torch.linalg.norm(r)) <= 1e-8 * bnorm
    # the oracle's rows of b - A x: a subset of a residual <= 1e-8 ||b||,
    # up to the FP64 difference of the two matvecs on those rows
    xs_h = res.x.cpu().numpy()
    ax = ora.matvec(xs_h)
    r_rows = ora.rhs() - ax
    assert np.linalg.norm(r_rows) <= 1e-8 * bnorm + 1e-10 * np.linalg.norm(ax)


def test_large_workload_rows_against_oracle_and_plain_solve():
    gpu, ora = _workload(4, 8, 64, with_bodies=True)
    _check_workload(gpu, ora)


def test_time_series_workload_rows_against_oracle_and_plain_solve():
    gpu, ora = _workload(3, 12, 96, with_bodies=False)
    _check_workload(gpu, ora)


def test_batched_preconditioner_all_2048_fibers():
    """Workload 2 as stated: 2048 fibers, N drawn from {16, 24, 32, 48, 64},
    A_sq from the fiber discretisation.  FP32 LU on MFMA + FP64 refinement:
    every block's relative residual <= 1e-12.  FP64 LU: pivots equal
    LAPACK's on 64 sampled blocks, factors within 1e-12."""
    cfg = configs.config(2)
    print("case batched preconditioner 2048 fibers", flush=True)
    fibers = configs.make_fibers(cfg.fibers, cfg.seed)
    S = FiberSystem(fibers, None, lu_precision="f32c", lu_refine=10, lu_tol=1e-14)
    assert S.bs.nb == 2048
    rng = np.random.default_rng(2)
    xs = torch.from_numpy(rng.normal(size=S.bs.nrows)).cuda()
    b = torch.empty_like(xs)
    S.gA(S.A_buf, xs, b)
    y = S.lu.solve(b)
    Ay = torch.empty_like(b)
    S.gA(S.A_buf, y, Ay)
    torch.cuda.synchronize()
    assert int(S.lu.info.max()) == 0
    r = (b - Ay).cpu().numpy()
    bh = b.cpu().numpy()
    worst = 0.0
    for k in range(S.bs.nb):
        o, n = S.bs.x_off_h[k], S.bs.sizes[k]
        worst = max(worst, np.linalg.norm(r[o:o + n]) / np.linalg.norm(bh[o:o + n]))
    print(f"worst block residual {worst:.2e}, max sweeps {int(S.lu.iters.max())}", flush=True)
    assert worst <= 1e-12
    lu64 = BatchedLU(S.bs, S.A_buf, precision="f64")
    torch.cuda.synchronize()
    assert int(lu64.info.max()) == 0
    piv = lu64.piv.cpu().numpy()
    for k in sorted(rng.choice(S.bs.nb, 64, replace=False).tolist()):
        A = _block(S, k)
        lu_ref, piv_ref = sla.lu_factor(A, check_finite=False)
        o, n = S.bs.x_off_h[k], S.bs.sizes[k]
        np.testing.assert_array_equal(piv[o:o + n], piv_ref)
        a = int(S.bs.a_off_h[k])
        LU = lu64.lu64[a:a + n * n].view(n, n).T.cpu().numpy()
        assert rel2(LU, lu_ref) < 1e-12, k


def test_raw_sweep_stresslet_65536():
    """Workload 1's stresslet half: 65536 points in the unit cube, unit
    normals, densities N(0,1) times weights U(0.5,1.5)/N, self term
    excluded.  FP32C against FP64 <= 1e-6 over all targets; FP64 against the
    oracle on 96 sampled targets <= 1e-10."""
    print("case raw sweep stresslet 65536", flush=True)
    rng = np.random.default_rng(1)
    N = 65536
    x = rng.uniform(0, 1, size=(N, 3))
    n = rng.normal(size=(N, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    q = rng.normal(size=(N, 3)) * (rng.uniform(0.5, 1.5, size=N) / N)[:, None]
    X = torch.from_numpy(x).cuda()
    Q = torch.from_numpy(q).cuda()
    NR = torch.from_numpy(n).cuda()
    u64 = FK.pairsum("stresslet", X, Q, X, src_nrm=NR, exclude_self=True, precision="f64")
    u32 = FK.pairsum("stresslet", X, Q, X, src_nrm=NR, exclude_self=True, precision="f32c")
    torch.cuda.synchronize()
    e = rel2(u32.cpu().numpy(), u64.cpu().numpy())
    print(f"stresslet f32c vs f64: {e:.3e}", flush=True)
    assert e < 1e-6
    rows = rng.choice(N, size=96, replace=False)
    ref = K.stresslet_kernel_source_target(x, x[rows], n, q, exclude=(rows, rows + 1))
    assert rel2(u64.cpu().numpy()[rows], ref.reshape(-1, 3)) < 1e-10


@pytest.fixture(scope="module")
def small_chebyshev():
    cfg = configs.config(0, discretisation="chebyshev")
    print("case small workload, Chebyshev-Lobatto", flush=True)
    fibers = configs.make_fibers(cfg.fibers, cfg.seed, discretisation="chebyshev")
    xs, ns, ws = configs.make_periphery(cfg.periphery_nodes, cfg.periphery_radius)
    kw = dict(eta=cfg.eta, dt=cfg.dt, E=cfg.fibers.E, epsilon=cfg.fibers.epsilon,
              discretisation="chebyshev")
    ora = OT.HydroSystem(fibers, (xs, ns, ws), **kw)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws, eta=cfg.eta), **kw)
    return ora, gpu


def test_chebyshev_assembly_rhs_and_precond(small_chebyshev):
    """A_sq against the oracle to 1e-10 or ten times the reference's own
    rounding floor (spectral D_k of 32 nodes reach 1e8, so the floor is
    ~1e-9 here); right-hand side and preconditioner to 1e-10."""
    ora, gpu = small_chebyshev
    blocks = gpu.dense_blocks_host()
    ks = list(range(len(blocks)))
    x = np.random.default_rng(3).normal(size=ora.size)
    fl_blk, _ = _floors(ora, ks, x, ora.matvec(x))
    for k in ks:
        assert rel2(blocks[k], ora.A_sq[k]) < _bar(fl_blk[k]), (k, fl_blk[k])
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-10
    x = np.random.default_rng(3).normal(size=ora.size)
    z = gpu.precond(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert rel2(z.cpu().numpy(), ora.precond(x)) < 1e-10


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32c", 1e-6)])
def test_chebyshev_matvec(small_chebyshev, precision, tol):
    """FP64: 1e-10 or ten times the reference's own rounding floor on this
    product; FP32-compensated: 1e-6."""
    ora, gpu = small_chebyshev
    x = np.random.default_rng(4).normal(size=ora.size)
    y = gpu.matvec(torch.from_numpy(x).cuda(), precision=precision)
    torch.cuda.synchronize()
    yo = ora.matvec(x)
    if precision == "f64":
        _, fl = _floors(ora, list(range(len(ora.fibers))), x, yo)
        tol = _bar(fl)
    e = rel2(y.cpu().numpy(), yo)
    print(f"chebyshev matvec {precision}: {e:.3e} (bar {tol:.2e})", flush=True)
    assert e < tol


def test_chebyshev_solve_to_1e8(small_chebyshev):
    """One plain GMRES solve to 1e-8: the GPU solution meets the oracle's
    system to the same tolerance, and agrees with the oracle's own GMRES
    solution to the conditioning of the tolerance."""
    ora, gpu = small_chebyshev
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-8, restart=150, maxiter=200)
    torch.cuda.synchronize()
    assert res.info == 0 and res.residual <= 1e-8
    x = res.x.cpu().numpy()
    r = np.linalg.norm(ora.rhs - ora.matvec(x)) / np.linalg.norm(ora.rhs)
    assert r <= 1.05e-8
    xo, info, _ = OT.gmres(ora.matvec, ora.rhs, ora.precond, tol=1e-10)
    assert info == 0
    assert rel2(x, xo) < 1e-5
