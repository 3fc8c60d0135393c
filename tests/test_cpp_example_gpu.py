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


EXE_DIEL = os.path.join(ROOT, "xfdtd_amd", "examples", "diel_sphere_scatter")
STEPS_DIEL = 300


def diel_example_scene():
    """examples/3d_diel_sphere_scatter.cpp:18-96 through the Python builder
    (Simulation{dl} takes the default cfl 0.99)."""
    dl = 7.5e-3
    size = 2 * 0.1 + 18 * dl
    sc = S.Scene(cell_size=dl, cfl=0.99)
    sc.add_object(S.Object("domain", S.Cube((-size / 2,) * 3, (size,) * 3), S.Material.air()))
    sc.add_object(S.Object("sphere", S.Sphere((0.0, 0.0, 0.0), 0.1), S.Material("sphere", 3.0, 2.0, 0.0, 0.0)))
    sc.boundaries = [S.PML(o, 8) for o in range(6)]
    lam = 20 * dl
    tau = lam / 6e8
    sc.tfsf = S.TFSF3D(18, 18, 18, 1.0, 0.0, 0.0, 0.0, S.GaussianWaveform(tau, 4.5 * tau, 1.0))
    sc.nffft = S.NffftFd(15, 15, 15, [1e9], [math.pi * i / 180 for i in range(181)], [0.0])
    return sc, tau


def test_cpp_dielectric_sphere_example_matches_oracle(tmp_path, oracle_threads):
    """The dielectric-sphere program: NF-FF output against the FP64 oracle
    (1e-10 of each quantity's peak), and incident_fft.npy against the DFT of
    the incident Gaussian sampled at t = j dt, j < T (tfsf.cpp:51-66 with
    util/dft.h): sum_j g(j dt) exp(-2 pi i f j dt) dt."""
    assert os.path.exists(EXE_DIEL), "build the library first (make -C xfdtd_amd/csrc)"
    out = tmp_path / "diel"
    print(f"running {EXE_DIEL} {out} {STEPS_DIEL} 64", flush=True)
    r = subprocess.run([EXE_DIEL, str(out), str(STEPS_DIEL), "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Dielectric Sphere Scatter" in r.stdout, r.stdout
    sc, tau = diel_example_scene()
    o = OracleSim(sc, STEPS_DIEL)
    o.step(STEPS_DIEL)
    ref = o.farfield()
    dt = o.dt
    o.close()
    for name in ("e_theta", "e_phi", "h_theta", "h_phi", "power_theta", "power_phi"):
        got = np.load(out / f"{name}.npy")
        want = ref[name]
        assert got.shape == want.shape, (name, got.shape, want.shape)
        scale = np.abs(want).max()
        assert scale > 0, name
        err = np.abs(got - want).max() / scale
        assert err <= 1e-10, (name, err)
    inc = np.load(out / "incident_fft.npy")
    t = np.arange(STEPS_DIEL) * dt
    g = np.exp(-(((t - 4.5 * tau) / tau) ** 2))
    want = np.array([np.sum(g * np.exp(-2j * math.pi * 1e9 * t)) * dt])
    assert inc.shape == (1,) and inc.dtype == np.complex128
    assert abs(inc[0] - want[0]) <= 1e-12 * abs(want[0]), (inc, want)


def test_cpp_host_monitor_hook(tmp_path):
    """xfdtd::HostMonitor (the reference's Monitor::init / update /
    outputData virtuals, monitor/monitor.h:28-37) on the host: the example's
    Ez recorder reads one cell through Simulation::getFieldBox after every
    step; its series equals the engine's point probe at that cell bit for
    bit, one value per step."""
    exe = os.path.join(ROOT, "xfdtd_amd", "examples", "host_monitor_probe")
    assert os.path.exists(exe), "build the library first (make -C xfdtd_amd/csrc)"
    out = tmp_path / "hm"
    steps = 150
    print(f"running {exe} {out} {steps} 64", flush=True)
    r = subprocess.run([exe, str(out), str(steps), "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    host, probe = np.load(out / "host_ez.npy"), np.load(out / "probe_ez.npy")
    assert host.shape == (steps,) and probe.shape == (steps,)
    assert np.abs(probe).max() > 0
    assert np.array_equal(host, probe)
