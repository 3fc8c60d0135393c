"""Field monitors (TimeDomainFieldMonitor, MovieMonitor) on the GPU: the
.npy frames a movie monitor writes while the run is stepped in uneven
chunks, and a field monitor's final slice, equal the oracle's fields over
the same grid boxes at the same steps, bit for bit in FP64 (the sample
schedule t % plot_step == 0 or t == T - 1, movie_monitor.cpp:46-59, and the
file names %04d.npy included)."""
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim
from test_parity_gpu import pec_scene
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _scene(tmp, T, plot):
    sc = pec_scene(n=30, pml=4, gap=3, radius_cells=5, incidence=(0.4, 0.7, 0.2))
    dl = sc.cell_size
    half = 15 * dl   # the grid spans [-half, half] on every axis
    # an XY plane of Ez through k = 14 over i in [3, 27), j in [5, 25)
    xy = S.TimeDomainFieldMonitor(S.Cube((-half + 3 * dl, -half + 5 * dl, -half + 14 * dl), (24 * dl, 20 * dl, 0.0)),
                                  S.XY, S.EZ, os.path.join(tmp, "movie"))
    sc.monitors.append(S.MovieMonitor(xy, T, plot))
    # a YZ plane of Hx through i = 9 over j in [2, 28), k in [4, 20), final only
    yz = S.TimeDomainFieldMonitor(S.Cube((-half + 9 * dl, -half + 2 * dl, -half + 4 * dl), (0.0, 26 * dl, 16 * dl)),
                                  S.YZ, S.HX, os.path.join(tmp, "final"), "hx_yz.npy")
    sc.monitors.append(yz)
    boxes = {"movie": (S.EZ, np.s_[3:27, 5:25, 14:15]), "final": (S.HX, np.s_[9:10, 2:28, 4:20])}
    return sc, boxes


@pytest.mark.parametrize("precision", [64, 32])
def test_movie_and_field_monitors_match_oracle(tmp_path, precision):
    from xfdtd_amd.sim import Simulation
    T, plot = 47, 7
    sc, boxes = _scene(str(tmp_path), T, plot)
    g = Simulation(sc, T, precision=precision)
    for n in (5, 13, 22, T - 40):
        g.step(n)
    g.output_data()
    frames = sorted(os.listdir(tmp_path / "movie"))
    due = [t for t in range(T) if t % plot == 0 or t == T - 1]
    assert frames == [f"{t:04d}.npy" for t in due]
    o = OracleSim(sc, T)
    comp, sl = boxes["movie"]
    for t in range(T):
        o.step(1)
        if t in due:
            got = np.load(tmp_path / "movie" / f"{t:04d}.npy")
            ref = o.get_field(comp)[sl]
            assert got.dtype == np.float64 and got.shape == ref.shape
            if precision == 64:
                assert np.array_equal(got, ref), t
            else:
                assert np.abs(got - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-30), t
    comp, sl = boxes["final"]
    got = np.load(tmp_path / "final" / "hx_yz.npy")
    ref = o.get_field(comp)[sl]
    assert got.shape == (1, 26, 16)
    if precision == 64:
        assert np.array_equal(got, ref)
    assert np.abs(ref).max() > 0
