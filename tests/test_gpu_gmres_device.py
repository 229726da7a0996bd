"""GMRES with the Arnoldi steps and the Hessenberg least-squares problem on
the device (libf3d f3d_gmres_*) against the host engine and the FP64 oracle:
same iterations and solution, a reused context, steps run ahead of the first
read, recorded graphs of whole Arnoldi steps, restarts, and the
updated-residual closing of an inexact inner operator, convergence on an
invariant subspace."""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd.gmres import GmresContext, gmres
from fiber3d_amd.system import FiberSystem, Shell
from oracle import tstep as OT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    print("case device gmres system", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(8, 16, 1.0, 2.0), seed=5)
    xs, ns, ws = configs.make_periphery(300, 3.5)
    S = FiberSystem(fibers, Shell(xs, ns, ws), lu_precision="f32c", lu_refine=0)
    return S, OT.HydroSystem(fibers, (xs, ns, ws))


def test_device_engine_matches_host_engine(system):
    S, ora = system
    kw = dict(tol=1e-10, restart=150, maxiter=200)
    dev = gmres(S.matvec, S.rhs, S.precond, engine="device", **kw)
    host = gmres(S.matvec, S.rhs, S.precond, engine="host", **kw)
    torch.cuda.synchronize()
    print("iterations device / host:", dev.iterations, host.iterations)
    assert dev.info == 0 and dev.residual <= 1e-10
    assert dev.iterations == host.iterations
    assert rel2(dev.x.cpu().numpy(), host.x.cpu().numpy()) < 1e-9
    x = dev.x.cpu().numpy()
    assert np.linalg.norm(ora.rhs - ora.matvec(x)) / np.linalg.norm(ora.rhs) <= 1.05e-10


def test_context_reuse_run_ahead_and_graphs(system):
    """One context for several solves; steps run ahead of the first read
    (the surplus ignored); whole Arnoldi steps replayed from graphs."""
    S, _ = system
    ctx = GmresContext(S.size, 150, S.dev)
    ref = gmres(S.matvec, S.rhs, S.precond, tol=1e-10, ctx=ctx)
    for expect in (0, 3, ref.iterations + 4):
        r = gmres(S.matvec, S.rhs, S.precond, tol=1e-10, ctx=ctx, expect=expect)
        assert r.iterations == ref.iterations
        assert rel2(r.x.cpu().numpy(), ref.x.cpu().numpy()) < 1e-13
    for _ in range(3):
        r = gmres(S.matvec, S.rhs, S.precond, tol=1e-10, ctx=ctx,
                  graph_keys={"matvec": "f64"})
        torch.cuda.synchronize()
        assert r.iterations == ref.iterations
        assert rel2(r.x.cpu().numpy(), ref.x.cpu().numpy()) < 1e-12
    assert ctx.captures == 1


def test_restarts_and_inner_operator(system):
    S, ora = system
    inner = lambda v, out=None: S.matvec(v, out=out, precision="f32c")
    r = gmres(S.matvec, S.rhs, S.precond, tol=1e-9, restart=2, maxiter=100)
    assert r.info == 0 and r.restarts > 1 and r.residual <= 1e-9
    h = gmres(S.matvec, S.rhs, S.precond, tol=1e-9, restart=2, maxiter=100, engine="host")
    assert r.iterations == h.iterations
    x0 = r.x * (1.0 + 1e-6)
    u = gmres(S.matvec, S.rhs, S.precond, tol=1e-9, x0=x0, inner=inner, update_residual=True,
              measure_inner=True, inner_rel_err=1e-6)
    torch.cuda.synchronize()
    assert u.info == 0 and u.guess_used and u.residual_bound <= 1e-9
    x = u.x.cpu().numpy()
    assert np.linalg.norm(ora.rhs - ora.matvec(x)) / np.linalg.norm(ora.rhs) <= 1.05e-9


def test_invariant_subspace_converges_in_three_steps():
    """A right-hand side in a 3-dimensional invariant subspace of A P^-1:
    the third step's residual estimate drops to rounding and the solution is
    exact."""
    n = 64
    d = torch.arange(1, n + 1, dtype=torch.float64, device="cuda")
    mv = lambda v, out=None: torch.mul(d, v, out=out) if out is not None else d * v
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    b[[3, 10, 40]] = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    r = gmres(mv, b, None, tol=1e-14)
    torch.cuda.synchronize()
    assert r.info == 0 and r.iterations == 3
    assert rel2((d * r.x).cpu().numpy(), b.cpu().numpy()) < 1e-13
