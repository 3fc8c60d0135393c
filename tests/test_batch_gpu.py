"""Batched stepping (xf_batch_*): several uneven simulations stepped together
must give exactly the bits each gives when stepped alone, in FP64 and FP32,
and one of them must match the FP64 oracle bit for bit.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(scene, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(scene, steps, precision=precision)


def scenes():
    """Three uneven problems: a PEC sphere at normal incidence, an oblique
    dielectric + lossy scene with NF-FF, and a current source with probes and
    kappa > 1 CPML (no TFSF)."""
    out = []
    dl = 5e-3
    out.append(S.cube_domain_scene(
        30, dl, 5, cfl=0.9, objects=[S.Object("pec", S.Sphere((0, 0, 0), 5 * dl), S.Material.pec())],
        tfsf_gap=9, nffft_gap=7, freqs=[S.C_0 / (20 * dl)], theta=np.linspace(0, math.pi, 7), phi=[0.0]))
    dl = 4e-3
    objs = [S.Object("diel", S.Sphere((0.004, -0.002, 0.0), 6 * dl), S.Material("d", 4.0, 1.0, 0.0, 0.0)),
            S.Object("lossy", S.Cube((-0.02, -0.02, -0.02), (0.012, 0.016, 0.008)),
                     S.Material("l", 2.5, 1.5, 0.05, 0.0))]
    out.append(S.cube_domain_scene(37, dl, 6, cfl=0.99, objects=objs, tfsf_gap=10, nffft_gap=8,
                                   freqs=[S.C_0 / (30 * dl), S.C_0 / (15 * dl)],
                                   theta=np.linspace(0, math.pi, 9), phi=[0.3], incidence=(0.6, 0.9, 0.3)))
    dl = 4e-3
    n, d = 26, 5
    sc = S.cube_domain_scene(n, dl, d, cfl=0.95, objects=[
        S.Object("m", S.Cube((-0.012, -0.012, -0.012), (0.024, 0.024, 0.024)), S.Material("m", 2.0, 1.0, 0.0, 0.0))])
    sc.boundaries = [S.PML(o, d, 3, 0.8, 0.0, 0.05, 4.0) for o in range(6)]
    dt = S.courant_dt(dl, 0.95)
    half = (n - 2 * d) * dl / 2 + d * dl
    x0, y0, z0 = -half + 9 * dl, -half + 13 * dl, -half + 13 * dl
    sc.sources.append(S.CurrentSource(S.Cube((x0, y0, z0), (dl, dl, dl)), S.ZP, math.inf,
                                      S.GaussianWaveform(15 * dt, 60 * dt, 1.0)))
    sc.probes = [S.Probe(S.EZ, 9, 13, 13), S.Probe(S.HX, 16, 12, 13)]
    out.append(sc)
    return out


TILINGS = [(64, 4, 8, 0), (256, 0, 5, 0), (128, 2, 16, 16)]


def _state(s):
    st = {f"f{c}": s.get_field(c) for c in range(6)}
    if s.scene.nffft is not None:
        st["nf"] = s.get_nffft()
    if s.scene.probes:
        st["probes"] = s.get_probes()
    return st


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_bitwise_equals_single(precision):
    from xfdtd_amd.sim import Batch
    steps = 70
    sc = scenes()
    alone = [_sim(x, steps, precision) for x in sc]
    for s, t in zip(alone, TILINGS):
        s.set_tiling(*t)
        s.step(steps)
    together = [_sim(x, steps, precision) for x in sc]
    for s, t in zip(together, TILINGS):
        s.set_tiling(*t)
    b = Batch(together)
    b.step(30)
    b.step(steps - 30)
    b.sync()
    for q, (a, c) in enumerate(zip(alone, together)):
        assert c.current_step == steps
        sa, sc_ = _state(a), _state(c)
        for k in sa:
            assert np.array_equal(sa[k], sc_[k]), (q, k)
        assert np.abs(sa["f2"]).max() > 0
    b.close()


def test_batch_member_matches_oracle(oracle_threads):
    from xfdtd_amd.sim import Batch
    steps = 60
    sc = scenes()
    together = [_sim(x, steps) for x in sc]
    b = Batch(together)
    b.step(steps)
    b.sync()
    o = OracleSim(sc[1], steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(together[1].get_field(c), o.get_field(c)), c
    assert np.array_equal(together[1].get_nffft(), o.get_nffft())
    b.close()


def test_engine_steps_after_batch_continue_exactly():
    """Batch steps, then the members' own steps: the same bits as one run."""
    from xfdtd_amd.sim import Batch
    steps = 50
    sc = scenes()[:2]
    ref = [_sim(x, steps) for x in sc]
    for s in ref:
        s.step(steps)
    mix = [_sim(x, steps) for x in sc]
    b = Batch(mix)
    b.step(20)
    for s in mix:
        s.step(steps - 20)
    for a, c in zip(ref, mix):
        for comp in range(6):
            assert np.array_equal(a.get_field(comp), c.get_field(comp))
    b.close()


def test_batch_refuses_bad_members():
    from xfdtd_amd.sim import Batch
    sc = scenes()
    a, c = _sim(sc[0], 10, 64), _sim(sc[1], 10, 32)
    with pytest.raises(RuntimeError, match="precision"):
        Batch([a, c])
    d = _sim(sc[1], 10, 64)
    d.step(1)
    with pytest.raises(RuntimeError, match="step"):
        Batch([a, d])
    e = _sim(sc[1], 10, 64)
    e.set_tiling(64, 2, 8, 0)
    with pytest.raises(RuntimeError, match="threads"):
        Batch([a, e])
    f = _sim(sc[1], 10, 64)
    b = Batch([a, f])
    with pytest.raises(RuntimeError, match="already"):
        Batch([f])
    with pytest.raises(RuntimeError, match="beyond"):
        b.step(11)
    b.close()


def test_batch_tiling_tuner_picks_a_candidate_and_keeps_results():
    """autotune.tune_batch times each candidate on the batched launch and
    returns one of them; a batch tiled with it steps bit-identically to the
    members stepped alone with their default tiling."""
    from xfdtd_amd import autotune
    from xfdtd_amd.sim import Batch
    scs = scenes()
    probe = [_sim(sc, 60, 32) for sc in scs]
    best = autotune.tune_batch(probe, candidates=autotune.BATCH_CANDIDATES[:3], steps=10, warmup=2)
    assert best in autotune.BATCH_CANDIDATES[:3]
    for s in probe:
        s.close()
    alone = [_sim(sc, 40, 32) for sc in scs]
    for s in alone:
        s.step(40)
    tuned = [_sim(sc, 40, 32) for sc in scs]
    for s in tuned:
        s.set_tiling(*best)
    b = Batch(tuned)
    b.step(40)
    b.sync()
    b.close()
    for a, t in zip(alone, tuned):
        sa, st = _state(a), _state(t)
        for k in sa:
            assert np.array_equal(sa[k], st[k]), k
