"""Recorded HIP graphs of the Arnoldi operator A P^-1 against eager calls, and
GMRES / time steps driven by them against the FP64 oracle."""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd.gmres import gmres
from fiber3d_amd.graphs import ArnoldiGraphs
from fiber3d_amd.system import FiberSystem, Shell
from fiber3d_amd.tstep import TimeStepper
from oracle import tstep as OT

pytestmark = pytest.mark.gpu


def _small(with_bodies=True, shell_pc="auto"):
    fibers = configs.make_fibers(configs.FiberSpec(6, 16, 1.0, 2.0), seed=9)
    bodies = configs.make_bodies(2, 0.5, 120, 2.2, fibers=fibers, seed=1) if with_bodies else None
    xs, ns, ws = configs.make_periphery(300, 4.0)
    return fibers, bodies, (xs, ns, ws), Shell(xs, ns, ws, pc=shell_pc, pc_f32=True)


@pytest.mark.parametrize("shell_pc", ["dense", "block"])
def test_graph_replays_match_eager(shell_pc):
    """Several replays, with eager work and synchronisations between them, each
    equal to the eager precond + matvec of the same vector."""
    print(f"case graph replay shell={shell_pc}", flush=True)
    fibers, bodies, _, shell = _small(shell_pc=shell_pc)
    S = FiberSystem(fibers, shell, bodies=bodies, lu_precision="f32c", lu_refine=0)
    G = ArnoldiGraphs(S)
    rng = np.random.default_rng(3)
    # FP32 first: recording the FP64 graph afterwards grows the shared pair-sum
    # workspace, and the FP32 graph must keep working on its old buffer
    ops = {p: G.get(p) for p in ("f32c", "f64")}
    for _ in range(4):
        for precision, op in ops.items():
            v = torch.from_numpy(rng.normal(size=S.size)).cuda()
            ref = S.matvec(S.precond(v), precision=precision)
            torch.cuda.synchronize()
            got = op(v)
            torch.cuda.synchronize()
            assert rel2(got.cpu().numpy(), ref.cpu().numpy()) < 1e-13, precision
    assert all(op.replays == 4 for op in ops.values())
    assert G.captures == 2


def test_graph_gmres_reaches_fp64_tolerance():
    """GMRES with recorded FP32-compensated and FP64 operators: the solution
    meets 1e-8 in the oracle's own FP64 system."""
    print("case graph gmres", flush=True)
    fibers, bodies, shell_t, shell = _small()
    force = np.array([[0.3, -0.2, 0.5, 0.1, 0.0, -0.2], [0.0, 0.4, -0.1, 0.0, 0.3, 0.1]])
    S = FiberSystem(fibers, shell, bodies=bodies, body_force=force, lu_precision="f32c",
                    lu_refine=0)
    G = ArnoldiGraphs(S)
    res = gmres(S.matvec, S.rhs, S.precond, tol=1e-8,
                inner=lambda v, out=None: S.matvec(v, out=out, precision="f32c"),
                fused_inner=G.get("f32c"), fused_matvec=G.get("f64"))
    torch.cuda.synchronize()
    assert res.info == 0 and res.residual <= 1e-8
    ora = OT.HydroSystem(fibers, shell_t, bodies=bodies, body_force=force)
    x = res.x.cpu().numpy()
    r = np.linalg.norm(ora.rhs - ora.matvec(x)) / np.linalg.norm(ora.rhs)
    assert r <= 1.05e-8


def test_graph_time_steps_match_eager_steps():
    """Fast-mode steps with recorded operators track the eager steps over
    several re-assemblies; every step's residual is the FP64 operator's."""
    print("case graph time steps", flush=True)
    fibers, _, _, shell = _small(with_bodies=False)
    runs = {}
    for graph in (False, True):
        ts = TimeStepper(fibers, shell, mode="fast", tol=1e-10, graph=graph)
        assert ts.sys.pair_precision == "f64"
        stats = [ts.step() for _ in range(3)]
        torch.cuda.synchronize()
        assert all(st.residual <= 1e-10 for st in stats)
        runs[graph] = ts.sys.X.cpu().numpy()
        if graph:
            # each Arnoldi step's graph (one per operator precision) is
            # recorded once and replayed by every later step (assembly refills
            # the recorded buffers in place)
            assert 1 <= ts.gctx.captures <= 2
    assert rel2(runs[True], runs[False]) < 1e-12


def test_graph_scratch_outlives_recording():
    """The recorded operator's scratch buffer stays allocated: tensors created
    after recording never share its memory (a replay writes it every call)."""
    print("case graph scratch lifetime", flush=True)
    fibers, bodies, _, shell = _small()
    S = FiberSystem(fibers, shell, bodies=bodies, lu_precision="f32c", lu_refine=0)
    op = ArnoldiGraphs(S).get("f32c")
    scratch = op.keep[0].data_ptr()
    fresh = [torch.zeros(S.size, dtype=torch.float64, device="cuda") for _ in range(64)]
    assert all(t.data_ptr() != scratch for t in fresh)
    for _ in range(3):
        op(torch.from_numpy(np.random.default_rng(1).normal(size=S.size)).cuda())
    torch.cuda.synchronize()
    assert all(float(t.abs().max()) == 0.0 for t in fresh)


def test_graph_of_attached_system_survives_steps():
    """Recorded Arnoldi steps of a system with fibers clamped to bodies stay
    equal to eager ones after steps re-assemble it (the clamped rows, the
    tangents and the links are refilled in place): the trajectories agree."""
    print("case graph attached", flush=True)
    from test_gpu_attached import _setup as attached_setup
    fibers, bodies, attached, (xs, ns, ws) = attached_setup()
    force = [np.array([0.5, 0.0, 0.0, 0.0, 0.0, 0.3]), np.zeros(6)]
    runs = {}
    for graph in (False, True):
        ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-10, bodies=bodies,
                         body_force=force, attached=attached, graph=graph)
        for _ in range(4):
            assert ts.step().residual <= 1e-10
        torch.cuda.synchronize()
        runs[graph] = ts.sys.X.cpu().numpy()
        if graph:
            assert 1 <= ts.gctx.captures <= 2
    # the body rows sum over surface nodes with atomic adds (index_add_), so
    # two runs of this system differ at rounding level whether recorded or
    # not; solves to 1e-10 then agree to about the conditioning times that
    assert rel2(runs[True], runs[False]) < 1e-7
