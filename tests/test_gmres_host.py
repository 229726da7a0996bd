"""GMRES host logic on CPU tensors: convergence against a direct solve,
restarts, the inexact inner operator and its accuracy-floor switch, initial
guesses (accepted, rejected) and the fused A P^-1 path."""
import numpy as np
import torch
from scipy.linalg import block_diag

from fiber3d_amd.gmres import gmres


def _system(n=160, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.normal(size=n)
    At = torch.from_numpy(A)
    mv = lambda v, out=None: torch.mv(At, v, out=out) if out is not None else torch.mv(At, v)
    # block-diagonal (4x4) preconditioner: exact inverse of A's diagonal blocks
    Pm = torch.from_numpy(block_diag(*[np.linalg.inv(A[i:i + 4, i:i + 4])
                                       for i in range(0, n, 4)]))

    def pc(v, out=None):
        y = torch.mv(Pm, v)
        if out is None:
            return y
        out.copy_(y)
        return out
    return A, b, mv, pc, torch.from_numpy(b)


def test_converges_to_direct_solve():
    A, b, mv, pc, bt = _system()
    res = gmres(mv, bt, pc, tol=1e-12)
    assert res.info == 0 and res.residual <= 1e-12
    x = np.linalg.solve(A, b)
    assert np.linalg.norm(res.x.numpy() - x) / np.linalg.norm(x) < 1e-10
    # true residual recomputed independently
    assert np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b) <= 1.01e-12


def test_restarts_still_converge():
    A, b, mv, pc, bt = _system(seed=1)
    res = gmres(mv, bt, pc, tol=1e-10, restart=3, maxiter=200)
    assert res.info == 0 and res.restarts > 1
    assert np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b) <= 1.01e-10


def test_inexact_inner_operator_reaches_fp64_tolerance():
    """An inner operator with 1e-7 relative error cannot by itself bring the
    true residual to 1e-10; GMRES detects the floor and finishes with the
    exact operator."""
    A, b, mv, pc, bt = _system(seed=2)
    rng = np.random.default_rng(9)
    E = torch.from_numpy(rng.normal(size=A.shape) * 1e-7 * np.abs(A).max())
    Ai = torch.from_numpy(A) + E

    def inner(v, out=None):
        y = torch.mv(Ai, v)
        if out is None:
            return y
        out.copy_(y)
        return out
    res = gmres(mv, bt, pc, tol=1e-10, inner=inner)
    assert res.info == 0 and res.residual <= 1e-10
    assert res.inner_cycles >= 1 and res.restarts > res.inner_cycles
    assert np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b) <= 1.01e-10


def test_initial_guess_accepted_and_rejected():
    A, b, mv, pc, bt = _system(seed=3)
    x = np.linalg.solve(A, b)
    cold = gmres(mv, bt, pc, tol=1e-10)
    warm = gmres(mv, bt, pc, tol=1e-10, x0=torch.from_numpy(x * (1 + 1e-6)))
    assert warm.guess_used and warm.iterations < cold.iterations
    bad = gmres(mv, bt, pc, tol=1e-10, x0=torch.full_like(bt, float("nan")))
    assert not bad.guess_used and bad.info == 0
    far = gmres(mv, bt, pc, tol=1e-10, x0=torch.from_numpy(-100.0 * x))
    assert not far.guess_used and far.info == 0      # residual above ||b||: zero guess
    assert np.allclose(bad.x.numpy(), cold.x.numpy(), rtol=0, atol=1e-12)


def test_fused_operator_path_matches_separate_calls():
    A, b, mv, pc, bt = _system(seed=4)
    calls = {"n": 0}

    def fused(v, out):
        calls["n"] += 1
        out.copy_(mv(pc(v)))
        return out
    r1 = gmres(mv, bt, pc, tol=1e-10)
    r2 = gmres(mv, bt, pc, tol=1e-10, fused_matvec=fused)
    assert calls["n"] == r2.iterations
    assert r1.iterations == r2.iterations
    assert np.abs(r1.x.numpy() - r2.x.numpy()).max() < 1e-13


def test_zero_rhs():
    A, b, mv, pc, bt = _system(seed=5)
    res = gmres(mv, torch.zeros_like(bt), pc, tol=1e-10)
    assert res.info == 0 and res.iterations == 0 and float(res.x.abs().max()) == 0.0


def _inexact(A, rel, seed):
    rng = np.random.default_rng(seed)
    Ai = torch.from_numpy(A + rng.normal(size=A.shape) * rel * np.abs(A).max())

    def inner(v, out=None):
        y = torch.mv(Ai, v)
        if out is None:
            return y
        out.copy_(y)
        return out
    return inner


def test_residual_update_from_a_close_guess_needs_one_fp64_product():
    """From a guess with a small residual, the FP64-anchored update replaces
    the closing FP64 product; its bound holds for the true residual."""
    A, b, mv, pc, bt = _system(seed=6)
    x = np.linalg.solve(A, b)
    calls = {"n": 0}

    def counted(v, out=None):
        calls["n"] += 1
        return mv(v, out=out)
    x0 = torch.from_numpy(x + 1e-6 * np.random.default_rng(1).normal(size=x.shape))
    res = gmres(counted, bt, pc, tol=1e-10, x0=x0, inner=_inexact(A, 1e-7, 10),
                update_residual=True)
    assert res.info == 0 and res.residual_kind == "update" and calls["n"] == 1
    true = np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b)
    assert true <= res.residual_bound <= 1e-10
    assert abs(true - res.residual) <= 1e-6 * 1e-4     # inner error times ||b - A x0|| / ||b||


def test_residual_update_falls_back_to_fp64_from_zero():
    """From a zero guess the correction is as large as b, the inner operator's
    error on it exceeds the tolerance, and the FP64 residual decides."""
    A, b, mv, pc, bt = _system(seed=7)
    res = gmres(mv, bt, pc, tol=1e-10, inner=_inexact(A, 1e-7, 11), update_residual=True)
    assert res.info == 0 and res.residual_kind == "fp64" and res.residual <= 1e-10
    assert np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b) <= 1.01e-10


def test_residual_update_rejected_when_bound_fails():
    """A large declared inner error makes the bound fail even from a close
    guess: the FP64 product is computed and the result stays correct."""
    A, b, mv, pc, bt = _system(seed=8)
    x = np.linalg.solve(A, b)
    x0 = torch.from_numpy(x + 1e-6 * np.random.default_rng(2).normal(size=x.shape))
    res = gmres(mv, bt, pc, tol=1e-10, x0=x0, inner=_inexact(A, 1e-7, 12),
                update_residual=True, inner_rel_err=1.0)
    assert res.info == 0 and res.residual_kind == "fp64"
    assert np.linalg.norm(b - A @ res.x.numpy()) / np.linalg.norm(b) <= 1.01e-10


def test_inner_error_measured_at_fp64_closing():
    """measure_inner: a cycle that closes with the FP64 residual reports the
    inner operator's error on its own correction z = x - x0,
    ||(A_inner - A) z|| / ||A_inner z||, recomputed here from the matrices."""
    A, b, mv, pc, bt = _system(seed=10)
    rng = np.random.default_rng(13)
    E = rng.normal(size=A.shape) * 1e-8 * np.abs(A).max()
    Ai = torch.from_numpy(A + E)
    inner = lambda v, out=None: torch.mv(Ai, v, out=out) if out is not None else torch.mv(Ai, v)
    x = np.linalg.solve(A, b)
    x0 = x + 1e-5 * rng.normal(size=x.shape)
    res = gmres(mv, bt, pc, tol=1e-10, x0=torch.from_numpy(x0), inner=inner,
                measure_inner=True)
    assert res.residual_kind == "fp64" and res.inner_cycles == 1 and res.restarts == 1
    z = res.x.numpy() - x0
    ref = np.linalg.norm(E @ z) / np.linalg.norm((A + E) @ z)
    assert abs(res.inner_err - ref) <= 1e-4 * ref
    # without measure_inner (and no update) nothing is measured
    assert gmres(mv, bt, pc, tol=1e-10, x0=torch.from_numpy(x0), inner=inner).inner_err is None


def test_target_below_tolerance():
    """Once iterating, GMRES continues to `target`; a guess whose residual
    is already within tol is accepted with no iteration."""
    A, b, mv, pc, bt = _system(seed=9)
    x = np.linalg.solve(A, b)
    plain = gmres(mv, bt, pc, tol=1e-8)
    deep = gmres(mv, bt, pc, tol=1e-8, target=1e-11)
    assert deep.residual <= 1e-11 < plain.residual <= 1e-8
    assert deep.iterations > plain.iterations
    near = torch.from_numpy(x * (1 + 1e-12))
    res = gmres(mv, bt, pc, tol=1e-8, target=1e-11, x0=near)
    assert res.iterations == 0 and res.guess_used and res.residual <= 1e-8
