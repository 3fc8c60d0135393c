"""The seeded workload builders: the vectorised splitmix64 draws equal the
scalar ones bit for bit, and every workload scene plans on the CPU."""
import numpy as np

from xfdtd_amd import scene as S
from xfdtd_amd.workloads import SplitMix64, adjoint_design_scene, batched_adjoint_set


def test_uniform_array_equals_scalar_draws():
    for seed in (0, 0xC4, 0xC5, 2**64 - 1):
        a, b = SplitMix64(seed), SplitMix64(seed)
        x = np.array([a.uniform(1.0, 4.0) for _ in range(777)])
        assert np.array_equal(x, b.uniform_array(777, 1.0, 4.0))
        assert a.state == b.state
        assert a.next_u64() == b.next_u64()


def test_splitmix64_known_answer():
    # splitmix64 from state 0: the published first outputs
    r = SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_batched_adjoint_set_plans():
    from xfdtd_amd.sim import Simulation
    scs = batched_adjoint_set(count=2)
    assert len({(sc.tfsf.theta, sc.tfsf.phi) for sc in scs}) == 2
    for sc in scs:
        assert sc.design.box == (32, 64, 32, 64, 32, 64)
        assert any(isinstance(o.material, S.DrudeMedium) for o in sc.objects)
        g = Simulation(sc, 10, plan_only=True)
        assert g.dims == (96, 96, 96)
    sc = adjoint_design_scene()
    assert np.asarray(sc.design.eps_r).min() >= 1.0 and np.asarray(sc.design.eps_r).max() < 4.0
