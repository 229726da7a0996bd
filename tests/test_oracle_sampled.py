"""Pins of the sampled-rows oracle (oracle/sampled.py) against the full
system oracle (oracle/tstep.py) at a size the full oracle handles: every
sampled row of A_sq, the right-hand side and the matvec equals the full
oracle's, with fibers, free rigid bodies (with an external force and torque)
and a periphery, and for both discretisations."""
import os

import numpy as np
import pytest

from conftest import rel2
from fiber3d_amd import configs
from oracle import build as obuild
from oracle import sampled as OS
from oracle import tstep as OT


@pytest.fixture(scope="module", autouse=True)
def _clib():
    if not os.path.exists(obuild.OUT):
        obuild.build(verbose=False)


@pytest.mark.parametrize("disc", ["reference", "chebyshev"])
def test_sampled_rows_equal_full_oracle(disc):
    spec = configs.FiberSpec(7, 16, 1.0, 2.5)
    fibers = configs.make_fibers(spec, seed=9, discretisation=disc)
    shell = configs.make_periphery(260, 4.5)
    bodies = configs.make_bodies(2, 0.5, 60, 2.5, fibers=fibers, seed=3)
    bf = np.array([[0.3, -0.2, 0.1, 0.05, 0.02, -0.04], [0.0, 0.1, -0.3, 0.0, 0.01, 0.02]])
    kw = dict(eta=0.9, dt=2e-3, E=0.02, discretisation=disc, bodies=bodies, body_force=bf)
    full = OT.HydroSystem(fibers, shell, **kw)
    rows_f, rows_s = [1, 4, 6], [0, 17, 131, 259]
    smp = OS.SampledHydro(fibers, shell, fiber_rows=rows_f, shell_rows=rows_s, **kw)
    idx = smp.row_index()
    assert smp.size == full.size
    for k in rows_f:
        assert rel2(smp.A_sq[k], full.A_sq[k]) < 1e-14
    assert rel2(smp.rhs(), full.rhs[idx]) < 1e-12
    rng = np.random.default_rng(4)
    for _ in range(2):
        x = rng.normal(size=full.size)
        assert rel2(smp.matvec(x), full.matvec(x)[idx]) < 1e-12


def test_row_index_covers_the_sampled_blocks():
    spec = configs.FiberSpec(3, 8, 1.0, 1.5)
    fibers = configs.make_fibers(spec, seed=2)
    shell = configs.make_periphery(50, 3.0)
    smp = OS.SampledHydro(fibers, shell, fiber_rows=[2], shell_rows=[5])
    idx = smp.row_index()
    assert idx.tolist() == [15, 16, 17] + list(range(150 + 64, 150 + 96))
