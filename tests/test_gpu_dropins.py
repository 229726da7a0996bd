"""The remaining drop-ins of the reference's boundary against the oracle:
dense Oseen tensors ('r' and 'xyz' formats, source-target), the dense
stresslet-times-normal matrix, n n^T, the Periphery class's
singularity-subtraction vectors, and the block-diagonal fiber LU builder."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import rel2
from fiber3d_amd import kernels as FK
from fiber3d_amd.dense import build_block_diagonal_lu_preconditioner_fiber
from fiber3d_amd.geometry import shape_gallery_sphere
from fiber3d_amd.periphery import Periphery
from oracle import kernels as K
from oracle import periphery as P

pytestmark = pytest.mark.gpu


def _pts(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    x[3] = x[7]                      # a coincident pair: regularised entries
    x[5] = x[6] + 3e-6               # a pair inside eps
    return x


@pytest.mark.parametrize("fmt", ["r", "xyz"])
def test_oseen_tensor_dropins(fmt):
    print(f"case oseen tensor {fmt}", flush=True)
    s, t = _pts(1, 50), _pts(2, 40)
    cast = (lambda a: a) if fmt == "r" else (lambda a: a.T.ravel())
    G = FK.oseen_tensor_source_target(cast(s), cast(t), eta=0.7, input_format=fmt)
    ref = K.oseen_tensor_source_target(cast(s), cast(t), eta=0.7, input_format=fmt)
    assert G.shape == (120, 150) and rel2(G, ref) < 1e-13
    G = FK.oseen_tensor(cast(s), input_format=fmt)
    assert rel2(G, K.oseen_tensor(cast(s), input_format=fmt)) < 1e-13


def test_dense_stresslet_and_complementary_dropins():
    print("case dense stresslet / complementary", flush=True)
    x = _pts(3, 60)
    n = np.random.default_rng(4).normal(size=(60, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    S = FK.stresslet_kernel_times_normal_numba(x, n, eta=1.3)
    assert rel2(S, K.stresslet_kernel_times_normal(x, n, eta=1.3)) < 1e-12
    Nk = FK.complementary_kernel(x, n)
    assert np.array_equal(Nk, K.complementary_kernel(x, n))


def test_periphery_singularity_subtraction_vectors():
    print("case periphery dropin", flush=True)
    x, nout = shape_gallery_sphere(400, 3.0)
    w = np.full(400, 4.0 * np.pi * 3.0**2 / 400)
    per = Periphery(np.zeros(3), None, x, -nout, w)
    per.get_singularity_subtraction_vectors(1.0)
    ex, ey, ez = P.singularity_subtraction_vectors(x, -nout, w)
    for got, ref in ((per.ex, ex), (per.ey, ey), (per.ez, ez)):
        assert got.shape == (1200,) and rel2(got, ref.ravel()) < 1e-10


def test_block_diagonal_lu_builder_matches_lapack():
    print("case lu builder dropin", flush=True)
    rng = np.random.default_rng(5)
    blocks = [rng.normal(size=(n, n)) + 4.0 * np.eye(n) for n in (64, 96, 17, 256)]
    LU_all, P_all = build_block_diagonal_lu_preconditioner_fiber(None, blocks)
    for A, LU, piv in zip(blocks, LU_all, P_all):
        lu_ref, piv_ref = sla.lu_factor(A, check_finite=False)
        np.testing.assert_array_equal(piv, piv_ref)
        assert rel2(LU, lu_ref) < 1e-12
        b = rng.normal(size=A.shape[0])
        y = sla.lu_solve((LU, piv), b)            # the reference's apply, unchanged
        assert rel2(A @ y, b) < 1e-12
