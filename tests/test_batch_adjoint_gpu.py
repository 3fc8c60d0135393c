"""Batched gradients (xf_batch_grad_eps): every member's J and dJ/d eps_r
equal its own xf_engine_grad_eps bit for bit (FP64 and FP32, uneven
checkpoint segments and none), one member matches the oracle's gradient,
the members are left at the end of the forward run, and the windowed
replays give the no-recompute gradient under mixed tile shapes."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd.workloads import batched_adjoint_set

pytestmark = pytest.mark.gpu


def small_set():
    """Three members of the batched-gradient workload at n = 34: Drude block,
    design cube, TFSF at per-member incidence, objective plane past both."""
    return batched_adjoint_set(count=3, seed=0xC5, n=34, pml=5, box=(12, 18), drude=(8, 12), tfsf_gap=7)


def _sim(sc, steps, precision):
    from xfdtd_amd.sim import Simulation
    return Simulation(sc, steps, precision=precision)


@pytest.mark.parametrize("precision,every", [(64, 7), (64, 0), (32, 9)])
def test_batch_grad_equals_single(precision, every):
    from xfdtd_amd.sim import Batch
    steps = 60
    scs = small_set()
    single = [_sim(sc, steps, precision).grad(checkpoint_every=every) for sc in scs]
    members = [_sim(sc, steps, precision) for sc in scs]
    b = Batch(members)
    got = b.grad(checkpoint_every=every)
    for q, ((J1, g1), (J2, g2)) in enumerate(zip(single, got)):
        assert J1 == J2, q
        assert np.array_equal(g1, g2), q
        assert np.abs(g1).max() > 0
    # members end at the forward run's last step, with its fields
    ref = _sim(scs[1], steps, precision)
    ref.step(steps)
    assert members[1].current_step == steps
    for c in range(6):
        assert np.array_equal(members[1].get_field(c), ref.get_field(c)), c
    b.close()


def test_batch_grad_member_matches_oracle(oracle_threads):
    from xfdtd_amd.sim import Batch
    steps = 60
    scs = small_set()
    members = [_sim(sc, steps, 64) for sc in scs]
    b = Batch(members)
    got = b.grad(checkpoint_every=11)
    Jo, go = OracleSim(scs[2], steps).grad()
    assert got[2][0] == pytest.approx(Jo, rel=1e-14)
    assert np.abs(got[2][1] - go).max() <= 1e-10 * np.abs(go).max()
    b.close()


def test_grad_independent_of_checkpoint_interval():
    """The full-size member (96^3, 1200 steps) gives the same J and gradient,
    bit for bit, with no recompute and with 172 or 12 checkpointed segments,
    alone and in a batch.  (A fill of the objective buffer on the null
    stream once landed after the copy of the final DFT on the engine's
    stream, zeroing J and the gradient at a checkpoint interval of 7.)"""
    from xfdtd_amd.sim import Batch
    steps = 1200
    scs = batched_adjoint_set(count=2)
    runs = {}
    for every in (0, 7, 100):
        s = _sim(scs[1], steps, 64)
        runs[every] = s.grad(checkpoint_every=every)
        s.close()
    J0, g0 = runs[0]
    assert J0 > 0 and np.abs(g0).max() > 0
    for every in (7, 100):
        assert runs[every][0] == J0, every
        assert np.array_equal(runs[every][1], g0), every
    members = [_sim(sc, steps, 64) for sc in scs]
    b = Batch(members)
    got = b.grad(checkpoint_every=7)
    b.close()
    assert got[1][0] == J0
    assert np.array_equal(got[1][1], g0)


def test_batch_replay_windows_mixed_tilings():
    """Batched replays launch only each member's tiles within reach of its
    design box: with members on different tile shapes and orders (column
    order with padding workgroups, x fastest, flat runs), the checkpointed
    batched gradient equals the no-recompute one bit for bit."""
    from xfdtd_amd.sim import Batch
    steps = 60
    scs = small_set()
    tilings = [(64, 4, 8, 16), (64, 4, 4, 4), (256, 0, 4, 8)]
    out = []
    for every in (7, 0):
        members = [_sim(sc, steps, 64) for sc in scs]
        for m, t in zip(members, tilings):
            m.set_tiling(*t)
        b = Batch(members)
        out.append(b.grad(checkpoint_every=every))
        b.close()
        for m in members:
            m.close()
    for (J1, g1), (J2, g2) in zip(*out):
        assert J1 == J2
        assert np.abs(g2).max() > 0
        assert np.array_equal(g1, g2)
