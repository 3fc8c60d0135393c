"""The C++ host API end to end.

xfdtd_amd/examples/pec_sphere_bistatic_rcs is the reference's bistatic-RCS
program (examples/3d_pec_sphere_bistatic_rcs.cpp:34-103) written against
include/xfdtd/xfdtd.hpp.  It runs on the GPU in FP64 and writes the NF-FF
.npy files of NffftFd::outputData (nffft_fd.cpp:579-634); they must match the
FP64 oracle's far field for the same scene to 1e-10 of each quantity's peak
(the GPU reduces the face sums in a different order).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "xfdtd_amd", "examples", "pec_sphere_bistatic_rcs")
STEPS = 200


def example_scene():
    """The same scene through the Python builder (cfl 1 as the example passes)."""
    dl = 7.5e-3
    sc = S.Scene(cell_size=dl, cfl=1.0)
    sc.add_object(S.Object("domain", S.Cube((-40 * dl,) * 3, (80 * dl,) * 3),
                           S.Material("air", 1.0, 1.0, 0.0, 0.0)))
    sc.add_object(S.Object("scatter", S.Sphere((0.0, 0.0, 0.0), 0.1),
                           S.Material("pec", 1.0, 1.0, 1e10, 0.0)))
    sc.boundaries = [S.PML(o, 12) for o in range(6)]
    lam = 20 * dl
    tau = lam / 6e8
    sc.tfsf = S.TFSF3D(27, 27, 27, 1.0, 0.0, 0.0, 0.0, S.GaussianWaveform(tau, 4.5 * tau, 1.0))
    sc.nffft = S.NffftFd(17, 17, 17, [1e9], [math.pi * i / 180 for i in range(180)], [0.0])
    return sc


def test_cpp_bistatic_example_matches_oracle(tmp_path, oracle_threads):
    assert os.path.exists(EXE), "build the library first (make -C xfdtd_amd/csrc)"
    out = tmp_path / "nffft"
    print(f"running {EXE} {out} {STEPS} 64", flush=True)
    r = subprocess.run([EXE, str(out), str(STEPS), "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "104 x 104 x 104" in r.stdout, r.stdout
    o = OracleSim(example_scene(), STEPS)
    o.step(STEPS)
    ref = o.farfield()
    for name in ("e_theta", "e_phi", "h_theta", "h_phi", "n_theta", "n_phi", "l_theta", "l_phi",
                 "power_theta", "power_phi"):
        got = np.load(out / f"{name}.npy")
        want = ref[name]
        assert got.shape == want.shape, (name, got.shape, want.shape)
        scale = np.abs(want).max()
        assert scale > 0, name
        err = np.abs(got - want).max() / scale
        assert err <= 1e-10, (name, err)
    assert np.array_equal(np.load(out / "theta.npy"), np.array([math.pi * i / 180 for i in range(180)]))
    assert np.array_equal(np.load(out / "frequencies.npy"), np.array([1e9]))
