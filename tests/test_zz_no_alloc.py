"""The allocation-free far-field entry point (xf_engine_farfield_sums_noalloc)
against xf_engine_farfield_sums on the same engine: bit-identical output on
two calls in a row, the second with fewer directions, in FP64 and FP32; a
request larger than the scratch block is refused before any GPU work."""
import ctypes as C
import math

import numpy as np
import pytest

from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu

XF_EINVAL = 1


def _scene():
    dl = 5e-3
    objs = [S.Object("pec", S.Sphere((0, 0, 0), 6 * dl), S.Material.pec())]
    return S.cube_domain_scene(
        34, dl, 5, cfl=0.9, objects=objs, tfsf_gap=9, nffft_gap=7,
        freqs=[S.C_0 / (40 * dl), S.C_0 / (30 * dl), S.C_0 / (20 * dl)],
        theta=np.linspace(0, math.pi, 19), phi=[0.0, math.pi / 2], incidence=(0.3, 0.2, 0.1))


def _sums(L, name, eng, dl, k, theta, phi):
    nf = len(k)
    out = np.full(nf * len(theta) * len(phi) * 24, np.nan)
    dp = C.POINTER(C.c_double)
    th, ph, kk = (np.ascontiguousarray(v, dtype=np.float64) for v in (theta, phi, k))
    rc = getattr(L, name)(eng, dl, kk.ctypes.data_as(dp), th.ctypes.data_as(dp), len(th),
                          ph.ctypes.data_as(dp), len(ph), out.ctypes.data_as(dp))
    return rc, out


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_far_field_equals_allocating_entry_point(precision):
    from xfdtd_amd._lib import lib
    from xfdtd_amd.sim import Simulation
    L = lib()
    sc = _scene()
    steps = 100
    print(f"no-alloc far field: 34^3 PEC sphere, {steps} FP{precision} steps", flush=True)
    sim = Simulation(sc, steps, precision=precision)
    sim.step(steps)
    sim.sync()
    eng = sim.engine
    dl = sc.cell_size
    k = [2 * math.pi * f / S.C_0 for f in sc.nffft.frequencies]
    calls = [(np.linspace(0, math.pi, 181), [0.0, math.pi / 2, 1.0]),   # 543 directions
             (np.linspace(0.1, 2.9, 7), [0.4])]                        # then 7
    for theta, phi in calls:
        rc_a, a = _sums(L, "xf_engine_farfield_sums", eng, dl, k, theta, phi)
        rc_b, b = _sums(L, "xf_engine_farfield_sums_noalloc", eng, dl, k, theta, phi)
        assert rc_a == 0, L.xf_last_error().decode()
        assert rc_b == 0, L.xf_last_error().decode()
        assert np.isfinite(b).all()
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b), float(np.abs(a - b).max())
    # 3 frequencies x 6000 x 40 directions need 138 MB: refused on the host
    rc, _ = _sums(L, "xf_engine_farfield_sums_noalloc", eng, dl, k, np.linspace(0, math.pi, 6000),
                  np.linspace(0, 2 * math.pi, 40))
    assert rc == XF_EINVAL
    assert "scratch" in L.xf_last_error().decode()
    # the engine still works after the refusal
    rc_b, b = _sums(L, "xf_engine_farfield_sums_noalloc", eng, dl, k, *calls[1])
    assert rc_b == 0
    sim.close()
