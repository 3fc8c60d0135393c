"""Fields above 4 GiB per component.  The curl kernels address a tile's
planes with 32-bit byte offsets from a 64-bit base at the tile's first
plane, so a component may exceed 4 GiB.

Check: a 1800 x 800 x 800 vacuum grid (1.2e9 padded cells per component:
9.6 GB in FP64, 4.8 GB in FP32) with a soft z current source at cell
(1700, 400, 400), far past the first 4 GiB of every component, against a
64^3 grid with the same source at its centre.  Away from the walls the
update is the same in both, and after 14 steps every non-zero value lies
within 28 cells of the source (each step reaches one cell further through E
and one through H), so the source's 60^3 neighbourhood must be
bit-identical: the FP64 run against the oracle on the small grid, the FP32
run against the GPU on the small grid.  dl = 2^-8 m keeps the cell
coordinates exact.
"""
import math

import numpy as np
import pytest

from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu

DL = 2.0 ** -8
STEPS = 14
HALF = 30


def box_scene(n, src):
    sc = S.Scene(cell_size=DL, cfl=0.9)
    sc.add_object(S.Object("box", S.Cube((0.0, 0.0, 0.0), tuple(m * DL for m in n)), S.Material.air()))
    dt = S.courant_dt(DL, 0.9)
    sc.sources.append(S.CurrentSource(S.Cube(tuple(c * DL for c in src), (DL, DL, DL)), S.ZP, math.inf,
                                      S.GaussianWaveform(4 * dt, 6 * dt, 1.0)))
    return sc


def _box(g, comp, lo):
    """comp over the 60^3 box from `lo` (xf_engine_get_box: only the box
    crosses to the host)."""
    import ctypes as C

    from xfdtd_amd._lib import lib
    L = lib()
    box = (C.c_int64 * 6)(lo[0], 2 * HALF, lo[1], 2 * HALF, lo[2], 2 * HALF)
    out = np.empty((2 * HALF,) * 3)
    assert L.xf_engine_get_box(g.engine, comp, box, out.ctypes.data_as(C.POINTER(C.c_double))) == 0, \
        L.xf_last_error().decode()
    return out


@pytest.mark.parametrize("precision", [64, 32])
def test_fields_beyond_4gib_per_component(precision, oracle_threads):
    from oracle.oracle import OracleSim
    from xfdtd_amd.sim import Simulation
    big_src, small_src = (1700, 400, 400), (32, 32, 32)
    small = box_scene((64, 64, 64), small_src)
    if precision == 64:
        ref_sim = OracleSim(small, STEPS)
    else:
        ref_sim = Simulation(small, STEPS, precision=32)
    ref_sim.step(STEPS)
    lo_s = [c - HALF for c in small_src]
    ref = [ref_sim.get_field(c)[lo_s[0]:lo_s[0] + 2 * HALF, lo_s[1]:lo_s[1] + 2 * HALF,
                                lo_s[2]:lo_s[2] + 2 * HALF] for c in range(6)]
    ref_sim.close()
    print(f"{precision}-bit 1800x800x800 grid, source at {big_src}", flush=True)
    g = Simulation(box_scene((1800, 800, 800), big_src), STEPS, precision=precision)
    assert g.dims == (1800, 800, 800)
    g.step(STEPS)
    g.sync()
    lo_b = [c - HALF for c in big_src]
    for c in range(6):
        got = _box(g, c, lo_b)
        assert np.abs(ref[c]).max() > 0, c
        assert np.array_equal(got, ref[c]), (c, float(np.abs(got - ref[c]).max()))
    g.close()
