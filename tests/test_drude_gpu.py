"""Drude cells on the GPU against the oracle (FP64 bit for bit): fields,
NF-FF accumulators and the polarisation currents J, in a scene with TFSF,
CPML, two Drude objects and a plain object covering part of one; the
filled-cavity pin's scene; batched stepping with a Drude member; and
checkpoints, which must carry J.  (Gradients with Drude cells are covered
with the other adjoint tests.)"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from test_drude_cpu import NU, OMEGA_P, _cavity, _fit_residual, _roots
from test_host_cpu import drude_scene
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(sc, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(sc, steps, precision=precision)


def _j_box(o, box, c):
    return o.get_drude_j(c)[box[0]:box[1], box[2]:box[3], box[4]:box[5]]


def test_drude_scene_bitwise(oracle_threads):
    steps = 80
    sc = drude_scene()
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    assert np.array_equal(g.get_nffft(), o.get_nffft())
    box = g.drude_box()
    for c in range(3):
        jg = g.get_drude_j(c)
        assert np.array_equal(jg, _j_box(o, box, c)), c
        assert np.abs(jg).max() > 0


def _mode(sim):
    """TM110 mode of the PEC box in Ez, zero on the walls."""
    nx, ny, _ = sim.dims
    i = np.arange(nx + 1)[:, None, None]
    j = np.arange(ny + 1)[None, :, None]
    ez = np.sin(math.pi * i / nx) * np.sin(math.pi * j / ny) * np.ones(sim.field_shape(S.EZ))
    ez[0], ez[nx], ez[:, 0], ez[:, ny] = 0, 0, 0, 0
    return ez


def test_drude_cavity_bitwise_and_fp32_dispersion():
    steps = 300
    sc = _cavity(S.DrudeMedium("plasma", 1.0, OMEGA_P, NU))
    sc.probes = [S.Probe(S.EZ, 4, 3, 1)]
    runs = {}
    for prec in (64, 32):
        g = _sim(sc, steps, prec)
        g.set_field(S.EZ, _mode(g))
        g.step(steps)
        runs[prec] = g
    o = OracleSim(sc, steps)
    o.set_field(S.EZ, _mode(o))
    o.step(steps)
    for c in range(6):
        assert np.array_equal(runs[64].get_field(c), o.get_field(c)), c
    pr32 = runs[32].get_probes()[0]
    roots = _roots(runs[32].dt, 1.0, OMEGA_P, NU)
    assert _fit_residual(pr32, roots) < 1e-5


def test_drude_fp32_close_to_oracle(oracle_threads):
    steps = 80
    sc = drude_scene()
    g, o = _sim(sc, steps, 32), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        ref = o.get_field(c)
        err = np.abs(g.get_field(c) - ref).max() / np.abs(ref).max()
        print(f"Drude FP32 component {c} error {err:.3e}", flush=True)
        assert err < 1e-5, (c, err)


def test_drude_batch_member_bitwise():
    from xfdtd_amd.sim import Batch
    steps = 50
    scs = [drude_scene(), drude_scene(n=28, pml=4)]
    alone = [_sim(x, steps) for x in scs]
    for s in alone:
        s.step(steps)
    together = [_sim(x, steps) for x in scs]
    b = Batch(together)
    b.step(steps)
    b.sync()
    for a, c in zip(alone, together):
        for comp in range(6):
            assert np.array_equal(a.get_field(comp), c.get_field(comp)), comp
        for comp in range(3):
            assert np.array_equal(a.get_drude_j(comp), c.get_drude_j(comp)), comp
    b.close()


def test_drude_checkpoint_carries_j():
    steps = 60
    g = _sim(drude_scene(), steps)
    g.step(30)
    g.checkpoint(0)
    g.step(30)
    first = [g.get_field(c) for c in range(6)] + [g.get_drude_j(c) for c in range(3)]
    g.restore(0)
    g.step(30)
    again = [g.get_field(c) for c in range(6)] + [g.get_drude_j(c) for c in range(3)]
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
