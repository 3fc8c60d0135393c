"""Lumped elements, plane objects, voltage/current monitors and S-parameter
networks on the GPU against the oracle.

- The reference's basic lumped-element circuit with a capacitor, an
  inductor, a current source and a later object over part of a plate: FP64
  fields and every monitor series bit-identical to the oracle; FP32 within
  1e-5.
- The inductor-filled cavity of the oracle pin: bit-identical.
- A batch with that circuit as a member steps it bit for bit, and a
  checkpoint restore continues bit for bit (the inductor currents are state).
- Monitor files are [time, value] rows; the S parameters of a two-port
  network equal util/dft.h and network/port.cpp applied to the oracle's
  monitor series (1e-12).
- Gradients refuse engines with inductors.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from test_host_cpu import lumped_scene
from test_lumped_cpu import DL, N3
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(sc, steps, precision=64, output_dir=None):
    from xfdtd_amd.sim import Simulation
    return Simulation(sc, steps, precision=precision, output_dir=output_dir)


def test_lumped_circuit_bitwise(oracle_threads):
    steps = 300
    sc = lumped_scene()
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    mg, mo = g.get_line_monitors(), o.get_line_monitors()
    assert mg.shape == mo.shape == (4, steps)
    assert np.array_equal(mg, mo)
    for q in range(4):
        assert 0 < np.abs(mg[q]).max() < 1e3, q  # excited, and the circuit stays bounded


def test_lumped_circuit_fp32_close(oracle_threads):
    steps = 300
    sc = lumped_scene()
    g, o = _sim(sc, steps, 32), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    mg, mo = g.get_line_monitors(), o.get_line_monitors()
    for q in range(4):
        err = np.abs(mg[q] - mo[q]).max() / np.abs(mo[q]).max()
        print(f"lumped FP32 monitor {q} error {err:.3e}", flush=True)
        assert err < 1e-5, (q, err)


def test_inductor_cavity_bitwise():
    sc = S.Scene(cell_size=DL, cfl=0.99)
    sc.add_object(S.Object("cav", S.Cube((0, 0, 0), tuple(x * DL for x in N3)), S.Material("air", 1, 1, 0, 0)))
    sc.sources = [S.Inductor(S.Cube((0, 0, 0), tuple(x * DL for x in N3)), S.AX_Z, 2.8e-10)]
    steps = 200
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    for s in (g, o):
        nx, ny, _ = s.dims
        i = np.arange(nx + 1)[:, None, None]
        j = np.arange(ny + 1)[None, :, None]
        s.set_field(S.EZ, np.sin(math.pi * i / nx) * np.sin(math.pi * j / ny) * np.ones(s.field_shape(S.EZ)))
        s.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c


def test_lumped_batch_member_and_checkpoint():
    from xfdtd_amd.sim import Batch
    steps = 120
    alone = _sim(lumped_scene(), steps)
    alone.step(steps)
    members = [_sim(lumped_scene(), steps), _sim(lumped_scene(S.SinusoidalWaveform(0.5, 3e10, 0.2)), steps)]
    b = Batch(members)
    b.step(steps)
    b.sync()
    for c in range(6):
        assert np.array_equal(members[0].get_field(c), alone.get_field(c)), c
    assert np.array_equal(members[0].get_line_monitors(), alone.get_line_monitors())
    b.close()
    g = _sim(lumped_scene(), steps)
    g.step(60)
    g.checkpoint(0)
    g.step(60)
    first = [g.get_field(c) for c in range(6)]
    g.restore(0)
    g.step(60)
    for c in range(6):
        assert np.array_equal(g.get_field(c), first[c]), c


def test_monitor_files_and_network(tmp_path, oracle_threads):
    steps = 400
    sc = lumped_scene()
    for m in sc.monitors:
        m.output_dir = str(tmp_path / "mon")
    v1, i1, i2, v2 = sc.monitors
    freqs = np.linspace(1e9, 8e10, 9)
    sc.network = S.Network([S.Port(1, 50 + 0j, True, v1, i1), S.Port(2, 75 + 5j, False, v2, i2)], list(freqs),
                           str(tmp_path / "net"))
    g = _sim(sc, steps, output_dir=str(tmp_path))
    g.step(steps)
    g.output_data()
    o = OracleSim(sc, steps)
    o.step(steps)
    mo = o.get_line_monitors()
    dt = g.dt
    for q, m in enumerate(sc.monitors):
        a = np.load(tmp_path / "mon" / m.output_file_name)
        assert a.shape == (2, steps)
        assert np.array_equal(a[1], mo[q])
        start, stop = (0.0, steps * dt) if isinstance(m, S.VoltageMonitor) else (0.5 * dt, (steps + 0.5) * dt)
        assert np.allclose(a[0], np.linspace(start, stop, steps), rtol=1e-14, atol=0)

    def dft(x):
        t = np.arange(len(x)) * dt
        return np.exp(-2j * np.pi * np.outer(freqs, t)) @ x * dt

    def ab(v, i, z):
        k = math.sqrt(z.real)
        return 0.5 * (dft(v) + z * dft(i)) / k, 0.5 * (dft(v) - np.conj(z) * dft(i)) / k

    a1, b1 = ab(mo[0], mo[1], 50 + 0j)
    a2, b2 = ab(mo[3], mo[2], 75 + 5j)
    expect = {11: b1 / a1, 21: b2 / a1}
    got = g.network_sparams()
    assert sorted(got) == [11, 21]
    for key, want in expect.items():
        assert np.abs(got[key] - want).max() <= 1e-12 * np.abs(want).max(), key
        f = np.load(tmp_path / "net" / f"s{key}.npy")
        assert f.shape == (2, len(freqs)) and f.dtype == np.complex128
        assert np.array_equal(f[0].real, freqs)
        assert np.array_equal(f[1], got[key])


def test_gradient_refuses_inductors():
    from adjoint_scenes import design_scene
    sc, _ = design_scene()
    dl = sc.cell_size
    corner = -13 * dl + 7 * dl  # cell (7, 7, 7): outside the CPML and the design box
    sc.sources = list(sc.sources) + [S.Inductor(S.Cube((corner,) * 3, (dl, dl, dl)), S.AX_Z, 1e-9)]
    with pytest.raises(RuntimeError, match="inductor"):
        _sim(sc, 40).grad(checkpoint_every=10)


def test_cpp_quarter_wave_transformer_example(tmp_path):
    """The reference's S-parameter example (examples/s_parameter_network.cpp)
    through the C++ API, FP64, 5000 steps: the monitor and S-parameter files
    (S from the files' V and I through util/dft.h and network/port.cpp to
    1e-10), and the transformer's physics: its 10 mm 70.7 ohm section on
    eps_r 4.6 is a quarter wave near 4 GHz, where |S11| has its minimum
    (below 0.35; 3.2 .. 4.8 GHz), and |S11| stays below 1 from 0.5 to 7 GHz."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "xfdtd_amd", "examples",
                       "quarter_wave_transformer")
    assert os.path.exists(exe), "build the library first (make -C xfdtd_amd/csrc)"
    out = tmp_path / "qwt"
    steps = 5000
    print(f"running {exe} {out} {steps} 64", flush=True)
    r = subprocess.run([exe, str(out), str(steps), "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    v1, c1 = np.load(out / "v1.npy"), np.load(out / "c1.npy")
    assert v1.shape == c1.shape == (2, steps)
    s11, s21 = np.load(out / "s11.npy"), np.load(out / "s21.npy")
    f = s11[0].real
    assert np.allclose(f, np.arange(2e7, 8e9, 2e7), rtol=1e-14, atol=0)
    dt = v1[0][-1] / steps  # the voltage time row ends at T dt
    t = np.arange(steps) * dt  # util/dft.h samples at j dt

    def dft(x):
        return np.exp(-2j * np.pi * np.outer(f, t)) @ x * dt

    z = 50.0
    a1 = 0.5 * (dft(v1[1]) + z * dft(c1[1])) / math.sqrt(z)
    b1 = 0.5 * (dft(v1[1]) - z * dft(c1[1])) / math.sqrt(z)
    assert np.abs(s11[1] - b1 / a1).max() <= 1e-10 * np.abs(b1 / a1).max()
    mag = np.abs(s11[1])
    band = (f > 0.5e9) & (f < 7e9)
    assert np.all(mag[band] < 1.0)
    q = np.argmin(np.where(band, mag, np.inf))
    print(f"|S11| minimum {mag[q]:.3f} at {f[q] / 1e9:.2f} GHz; |S21| there {abs(s21[1][q]):.3f}", flush=True)
    assert 3.2e9 < f[q] < 4.8e9 and mag[q] < 0.35
