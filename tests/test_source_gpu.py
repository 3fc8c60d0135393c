"""Lumped current sources on the GPU: every orientation, series and parallel
source boxes and a resistive source give the oracle's fields bit for bit
(FP64), and the GPU fields satisfy the same charge balance the oracle is
pinned by."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from test_source_cpu import _box_scene, _node_charge, _soft_charge
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("orientation,cells,resistance", [
    (S.ZP, (1, 1, 1), math.inf), (S.XN, (1, 1, 1), math.inf), (S.YP, (1, 1, 1), math.inf),
    (S.ZP, (1, 1, 3), math.inf), (S.XN, (1, 2, 2), math.inf), (S.YN, (2, 1, 1), 50.0),
])
def test_source_bitwise_and_charge(orientation, cells, resistance):
    from xfdtd_amd.sim import Simulation
    steps = 45
    sc, wave, dt = _box_scene(orientation, cells, resistance)
    g = Simulation(sc, steps, precision=64)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    if resistance == math.inf:
        q = _node_charge(g)
        expect = _soft_charge(wave, dt, steps)
        assert abs(np.abs(q).sum() - 2 * expect) <= 1e-9 * expect
        assert abs(q.sum()) <= 1e-9 * expect
