"""Divergence detection (xf_engine_set_divergence_check): a run whose time
step is above the Courant limit grows without bound until its fields
overflow; the H kernels flag the first non-finite step and the engine stops
there with XF_ESTATE instead of stepping on through NaNs.

- cfl 1.3 from seeded random fields: step() raises within one check interval
  of the first non-finite H value, the reported step is the first step at
  which the oracle's H is non-finite, and further steps are refused until a
  checkpoint is restored;
- a stable run (cfl 0.9) checked every step never raises;
- in a batch the diverging member is named;
- the same in the fused step (scheme 1).
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def unstable(cfl=1.3, n=20):
    """An n^3 vacuum box with PEC walls (no CPML)."""
    sc = S.cube_domain_scene(n, 5e-3, 0, cfl=cfl)
    sc.boundaries = []
    return sc


def _seed(g, seed=11):
    rng = np.random.default_rng(seed)
    for c in range(6):
        g.set_field(c, rng.standard_normal(g.field_shape(c)))


def _first_nonfinite_step(sc, steps, precision):
    """Step index (0-based) whose H update first leaves a non-finite value,
    from the oracle stepped one step at a time (FP64), or from the GPU run
    itself for FP32."""
    from xfdtd_amd.sim import Simulation
    if precision == 64:
        o = OracleSim(sc, steps)
        _seed(o)
        for t in range(steps):
            o.step(1)
            if not all(np.isfinite(o.get_field(c)).all() for c in (3, 4, 5)):
                o.close()
                return t
        o.close()
        return None
    g = Simulation(sc, steps, precision=precision)
    g.set_divergence_check(0)
    _seed(g)
    for t in range(steps):
        g.step(1)
        if not all(np.isfinite(g.get_field(c)).all() for c in (3, 4, 5)):
            g.close()
            return t
    g.close()
    return None


@pytest.mark.parametrize("precision", [64, 32])
def test_unstable_run_stops_at_first_nonfinite_step(precision, oracle_threads):
    from xfdtd_amd.sim import Simulation
    sc = unstable()
    steps = 3000
    first = _first_nonfinite_step(sc, steps, precision)
    print(f"cfl 1.3 {precision}-bit: first non-finite H at step {first}", flush=True)
    assert first is not None and first > 10
    g = Simulation(sc, steps, precision=precision)
    every = 16
    g.set_divergence_check(every)
    g.checkpoint(0)
    _seed(g)
    with pytest.raises(RuntimeError, match=f"non-finite H values at step {first}"):
        g.step(steps)
    assert g.diverged_step == first
    # detected at the first check after it
    assert g.current_step == (first // every + 1) * every
    with pytest.raises(RuntimeError, match="diverged"):
        g.step(1)
    g.restore(0)
    assert g.diverged_step is None
    g.step(5)
    assert g.current_step == 5


def test_stable_run_checked_every_step():
    from xfdtd_amd.sim import Simulation
    sc = unstable(cfl=0.9)
    g = Simulation(sc, 400)
    g.set_divergence_check(1)
    _seed(g)
    g.step(400)
    assert g.diverged_step is None


def test_batch_names_diverging_member():
    from xfdtd_amd.sim import Batch, Simulation
    sims = [Simulation(unstable(cfl=0.9), 3000, precision=32), Simulation(unstable(cfl=1.3), 3000, precision=32)]
    for s in sims:
        s.set_divergence_check(50)
        _seed(s)
    b = Batch(sims)
    with pytest.raises(RuntimeError, match="batch member 1: non-finite"):
        b.step(3000)
    assert sims[0].diverged_step is None and sims[1].diverged_step is not None
    b.close()
    for s in sims:
        s.close()


def test_fused_step_flags_divergence():
    from xfdtd_amd.sim import Simulation
    sc = unstable()
    first = _first_nonfinite_step(sc, 3000, 32)
    g = Simulation(sc, 3000, precision=32)
    g.set_sweep(1)
    assert g.sweep_active
    g.set_divergence_check(8)
    _seed(g)
    with pytest.raises(RuntimeError, match=f"non-finite H values at step {first}"):
        g.step(3000)
