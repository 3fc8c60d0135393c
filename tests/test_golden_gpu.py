"""The GPU engine (FP64) against the golden fixtures (tests/golden/*.npz,
oracle outputs frozen by tests/golden/make_golden.py): fields, NF-FF
accumulators, broadband series, line monitors and polarisation currents
bit for bit, without running the oracle."""
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

from make_golden import scenes  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(scenes()))
def test_engine_matches_fixture(name):
    from xfdtd_amd.sim import Simulation
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    build, steps = scenes()[name]
    assert int(ref["steps"]) == steps
    g = Simulation(build(), steps, precision=64)
    assert g.dt == float(ref["dt"]) and tuple(g.dims) == tuple(ref["dims"])
    g.step(steps)
    for c, n in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
        assert np.array_equal(g.get_field(c), ref[n]), (name, n)
    if "nffft" in ref.files:
        assert np.array_equal(g.get_nffft(), ref["nffft"])
    if "bb_samples" in ref.files:
        assert np.array_equal(g.get_bb_samples(), ref["bb_samples"])
    if "line_monitors" in ref.files:
        assert np.array_equal(g.get_line_monitors(), ref["line_monitors"])
    if "jx" in ref.files:
        b = g.drude_box()
        for c, n in enumerate(("jx", "jy", "jz")):
            assert np.array_equal(g.get_drude_j(c), ref[n][b[0]:b[1], b[2]:b[3], b[4]:b[5]]), (name, n)
