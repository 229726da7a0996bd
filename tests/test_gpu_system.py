"""GPU fiber + shell system against the FP64 oracle (tstep.time_step_hydro path)."""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd.gmres import gmres
from fiber3d_amd.system import FiberSystem, Shell
from oracle import tstep as OT

pytestmark = pytest.mark.gpu


def _build(nfib, N, nper, R_ball, R_shell, seed, **kw):
    fibers = configs.make_fibers(configs.FiberSpec(nfib, N, 1.0, R_ball), seed=seed)
    xs, ns, ws = configs.make_periphery(nper, R_shell)
    ora = OT.HydroSystem(fibers, (xs, ns, ws), eta=1.0, dt=1e-3, E=0.01)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), eta=1.0, dt=1e-3, E=0.01, **kw)
    return ora, gpu


@pytest.fixture(scope="module")
def small():
    print("case system small (8 fibers x 16, shell 300)", flush=True)
    return _build(8, 16, 300, 2.0, 3.5, 5)


def test_assembly_and_rhs(small):
    ora, gpu = small
    torch.cuda.synchronize()
    blocks = gpu.dense_blocks_host()
    for k in range(len(blocks)):
        # straight fibers: xsss = D3 x is rounding noise, so two FP64 assemblies
        # differ by ~1e-11 in the tension rows; the FP64 bar is 1e-10
        assert rel2(blocks[k], ora.A_sq[k]) < 1e-10
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-12


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32c", 1e-6)])
def test_matvec(small, precision, tol):
    ora, gpu = small
    rng = np.random.default_rng(3)
    x = rng.normal(size=ora.size)
    y = gpu.matvec(torch.from_numpy(x).cuda(), precision=precision)
    torch.cuda.synchronize()
    assert rel2(y.cpu().numpy(), ora.matvec(x)) < tol


def test_precond(small):
    ora, gpu = small
    rng = np.random.default_rng(4)
    x = rng.normal(size=ora.size)
    y = gpu.precond(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert rel2(y.cpu().numpy(), ora.precond(x)) < 1e-10


def test_gmres_small_matches_dense_solve(small):
    ora, gpu = small
    xd = np.linalg.solve(ora.dense_operator(), ora.rhs)
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-10)
    torch.cuda.synchronize()
    assert res.info == 0
    assert rel2(res.x.cpu().numpy(), xd) < 1e-7


def test_config0_gmres_to_1e8():
    """Workload 0: 32 fibers x 32 nodes, 1024-node shell, one GMRES solve to 1e-8."""
    print("case config0", flush=True)
    cfg = configs.config(0)
    fibers = configs.make_fibers(cfg.fibers, cfg.seed)
    xs, ns, ws = configs.make_periphery(cfg.periphery_nodes, cfg.periphery_radius)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), eta=cfg.eta, dt=cfg.dt, E=cfg.fibers.E)
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-8, restart=150, maxiter=200)
    torch.cuda.synchronize()
    assert res.info == 0 and res.residual <= 1e-8
    # the GPU solution satisfies the oracle's own system to the same tolerance
    ora = OT.HydroSystem(fibers, (xs, ns, ws), eta=cfg.eta, dt=cfg.dt, E=cfg.fibers.E)
    x = res.x.cpu().numpy()
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-12
    r = np.linalg.norm(ora.rhs - ora.matvec(x)) / np.linalg.norm(ora.rhs)
    assert r <= 1.05e-8
    # the FP32-compensated inner operator reaches the same true FP64 residual
    res32 = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-8,
                  inner=lambda v, out=None: gpu.matvec(v, out=out, precision="f32c"))
    torch.cuda.synchronize()
    assert res32.info == 0 and res32.residual <= 1e-8
    assert rel2(res32.x.cpu().numpy(), x) < 1e-6


def test_block_jacobi_shell_preconditioner_same_solution():
    """The 3x3 block-Jacobi shell preconditioner (used for large shells) changes
    only the iteration count, not the solution."""
    print("case shell block pc", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(8, 16, 1.0, 2.0), seed=5)
    xs, ns, ws = configs.make_periphery(300, 3.5)
    dense = FiberSystem(fibers, Shell(xs, ns, ws, pc="dense"))
    block = FiberSystem(fibers, Shell(xs, ns, ws, pc="block"))
    r1 = gmres(dense.matvec, dense.rhs, dense.precond, tol=1e-10)
    r2 = gmres(block.matvec, block.rhs, block.precond, tol=1e-10)
    torch.cuda.synchronize()
    print("iterations dense / block:", r1.iterations, r2.iterations)
    assert r1.info == 0 and r2.info == 0
    assert rel2(r2.x.cpu().numpy(), r1.x.cpu().numpy()) < 1e-7


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32c", 1e-6)])
def test_bodies_matvec_precond_and_solve(precision, tol):
    """Fibers + 2 free rigid spheres + shell against the oracle."""
    print(f"case bodies {precision}", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(6, 16, 1.0, 2.0), seed=9)
    bodies = configs.make_bodies(2, 0.5, 120, 2.2, fibers=fibers, seed=1)
    force = np.array([[0.3, -0.2, 0.5, 0.1, 0.0, -0.2], [0.0, 0.4, -0.1, 0.0, 0.3, 0.1]])
    xs, ns, ws = configs.make_periphery(300, 4.0)
    ora = OT.HydroSystem(fibers, (xs, ns, ws), bodies=bodies, body_force=force)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), bodies=bodies, body_force=force,
                      pair_precision=precision)
    torch.cuda.synchronize()
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-12
    rng = np.random.default_rng(6)
    x = rng.normal(size=ora.size)
    y = gpu.matvec(torch.from_numpy(x).cuda())
    z = gpu.precond(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert rel2(y.cpu().numpy(), ora.matvec(x)) < tol
    assert rel2(z.cpu().numpy(), ora.precond(x)) < 1e-10
    if precision == "f64":
        res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-10)
        xo, info, _ = OT.gmres(ora.matvec, ora.rhs, ora.precond, tol=1e-10)
        torch.cuda.synchronize()
        assert res.info == 0 and info == 0
        assert rel2(res.x.cpu().numpy(), xo) < 1e-7


def test_bodies_precond_after_rigid_motion():
    """The body inverse built in the initial pose and rotated with the bodies
    equals the preconditioner the oracle factors afresh at the moved pose."""
    print("case bodies moved", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(4, 16, 1.0, 2.0), seed=3)
    bodies = configs.make_bodies(3, 0.5, 120, 2.2, fibers=fibers, seed=4)
    xs, ns, ws = configs.make_periphery(300, 4.0)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), bodies=bodies)
    rng = np.random.default_rng(11)
    for _ in range(2):
        vel = torch.from_numpy(rng.normal(size=(3, 6))).cuda()
        gpu.bodies.move(vel, 0.25)
    B = gpu.bodies
    nb = B.nb
    moved = [(B.centres[k].cpu().numpy(), B.pos[k * nb:(k + 1) * nb].cpu().numpy(),
              B.nrm[k * nb:(k + 1) * nb].cpu().numpy(), B.w[k * nb:(k + 1) * nb].cpu().numpy())
             for k in range(B.count)]
    ora = OT.HydroSystem(fibers, (xs, ns, ws), bodies=moved)
    x = rng.normal(size=ora.size)
    z = gpu.precond(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert rel2(z.cpu().numpy(), ora.precond(x)) < 1e-10


@pytest.mark.parametrize("N", [[16, 24, 32, 48, 64], 32])
def test_hip_assembly_matches_torch_assembly(N):
    """f3d_fiber_assemble (one launch per node count) reproduces the batched
    torch assembly: A_sq, force rows, tension projections, tangents, weights
    and the right-hand side, for mixed node counts and unequal lengths."""
    print(f"case hip assembly N={N}", flush=True)
    spec = configs.FiberSpec(40, N, 1.0, 4.0)
    fibers = configs.make_fibers(spec, seed=12)
    lengths = np.random.default_rng(3).uniform(0.8, 1.6, len(fibers))
    xs, ns, ws = configs.make_periphery(400, 6.0)
    shell = Shell(xs, ns, ws)
    ref = FiberSystem(fibers, shell, lengths=lengths, assembly="torch")
    hip = FiberSystem(fibers, shell, lengths=lengths, assembly="hip")
    ora = OT.HydroSystem(fibers, (xs, ns, ws), lengths=lengths)
    torch.cuda.synchronize()
    # A_sq carries (2/L)^4 D4 entries up to ~1e9 that cancel (N = 64, L = 0.8):
    # FP64 assemblies in different summation orders differ by up to ~2e-10 per
    # block. The HIP assembly must be no further from the oracle than the torch
    # assembly (pinned at 1e-10 on the small system).
    bh, bt = hip.dense_blocks_host(), ref.dense_blocks_host()
    for k in range(len(bh)):
        e_h, e_t = rel2(bh[k], ora.A_sq[k]), rel2(bt[k], ora.A_sq[k])
        assert e_h < 1e-9 and e_h <= 3.0 * e_t + 1e-12, (k, e_h, e_t)
    assert rel2(hip.A_buf.cpu().numpy(), ref.A_buf.cpu().numpy()) < 1e-9
    assert rel2(hip.F_buf.cpu().numpy(), ref.F_buf.cpu().numpy()) < 1e-12
    for c in range(3):
        assert rel2(hip.Q_buf[c].cpu().numpy(), ref.Q_buf[c].cpu().numpy()) < 1e-12
    assert rel2(hip.xs.cpu().numpy(), ref.xs.cpu().numpy()) < 1e-12
    assert rel2(hip.weights.cpu().numpy(), ref.weights.cpu().numpy()) < 1e-15
    assert rel2(hip.rhs.cpu().numpy(), ora.rhs) < 1e-12
    x = np.random.default_rng(4).normal(size=ref.size)
    y = hip.matvec(torch.from_numpy(x).cuda())
    yt = ref.matvec(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    yo = ora.matvec(x)
    assert rel2(y.cpu().numpy(), yo) <= 3.0 * rel2(yt.cpu().numpy(), yo) + 1e-12
    assert rel2(y.cpu().numpy(), yo) < 1e-9


def test_ellipsoidal_periphery_matches_oracle():
    """An ellipsoidal shell (semi-axes 3, 2.5, 2; 900 nodes) around 12
    fibers: right-hand side, operator in FP64 and FP32-compensated,
    preconditioner and a GMRES solve against the oracle."""
    print("case ellipsoidal periphery", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(12, 16, 0.6, 1.0), seed=5)
    shell = configs.make_periphery_ellipsoid(900, 3.0, 2.5, 2.0)
    L = np.full(len(fibers), 0.6)
    gpu = FiberSystem(fibers, Shell(*shell), lengths=L)
    ora = OT.HydroSystem(fibers, shell, lengths=L)
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-11
    x = np.random.default_rng(8).normal(size=ora.size)
    xt = torch.from_numpy(x).cuda()
    ref = ora.matvec(x)
    assert rel2(gpu.matvec(xt).cpu().numpy(), ref) < 1e-10
    assert rel2(gpu.matvec(xt, precision="f32c").cpu().numpy(), ref) < 1e-6
    assert rel2(gpu.precond(xt).cpu().numpy(), ora.precond(x)) < 1e-9
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-10)
    torch.cuda.synchronize()
    assert res.info == 0
    xo = res.x.cpu().numpy()
    assert np.linalg.norm(ora.rhs - ora.matvec(xo)) / np.linalg.norm(ora.rhs) < 1.05e-10
