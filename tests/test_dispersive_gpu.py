"""Debye and Lorentz media on the GPU against the oracle.

- FP64 bit for bit: fields, NF-FF accumulators and J in a scene with TFSF,
  CPML, a Lorentz sphere, a lossy Debye cube, a Drude cube and a plain
  object covering part of the sphere; the filled TM110 cavities of the
  oracle pins.
- FP32: the Lorentz cavity's probe within 1e-5 of the oracle's peak.
- Checkpoints carry the Lorentz state (J, and the previous step's J and
  E): a restored run continues bit for bit.
- Gradients refuse Debye and Lorentz media.
- The reference's Lorentz-sphere monostatic test through the C++ API
  against the Mie series of the dispersive sphere.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from test_dispersive_cpu import DEBYE, LORENTZ, _medium
from test_drude_cpu import _cavity
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(sc, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(sc, steps, precision=precision)


def mixed_scene(n=30, dl=4e-3, pml=5):
    wp = 2 * math.pi * 20e9
    objs = [S.Object("lorentz", S.Sphere((0.002, -0.004, 0.0), 7 * dl),
                     S.LorentzMedium("l", 3.0, 1.5, wp, 2 * math.pi * 2e9)),
            S.Object("debye", S.Cube((-0.036, 0.01, -0.02), (0.02, 0.012, 0.016)),
                     S.DebyMedium("d", 5.0, 2.0, 1 / (2 * math.pi * 8e9), 0.03, 0.0)),
            S.Object("drude", S.Cube((0.02, 0.012, -0.03), (0.012, 0.012, 0.012)),
                     S.DrudeMedium("p", 1.0, 0.7 * wp, 2 * math.pi * 1e9)),
            S.Object("cap", S.Cube((0.004, -0.03, -0.03), (0.03, 0.06, 0.06)), S.Material("c", 3.0, 1.2, 0.01, 0.0))]
    return S.cube_domain_scene(n, dl, pml, cfl=0.99, objects=objs, tfsf_gap=pml + 2, nffft_gap=pml + 1,
                               freqs=[S.C_0 / (20 * dl)], theta=[0.0, 1.0], phi=[0.0])


def test_mixed_dispersive_scene_bitwise(oracle_threads):
    steps = 80
    sc = mixed_scene()
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    assert np.array_equal(g.get_nffft(), o.get_nffft())
    b = g.drude_box()
    for c in range(3):
        jg = g.get_drude_j(c)
        assert np.array_equal(jg, o.get_drude_j(c)[b[0]:b[1], b[2]:b[3], b[4]:b[5]]), c
        assert np.abs(jg).max() > 0


def _mode(sim):
    nx, ny, _ = sim.dims
    i = np.arange(nx + 1)[:, None, None]
    j = np.arange(ny + 1)[None, :, None]
    ez = np.sin(math.pi * i / nx) * np.sin(math.pi * j / ny) * np.ones(sim.field_shape(S.EZ))
    ez[0], ez[nx], ez[:, 0], ez[:, ny] = 0, 0, 0, 0
    return ez


@pytest.mark.parametrize("kind,p", [("debye", DEBYE[1]), ("lorentz", LORENTZ[0]), ("lorentz", LORENTZ[1])])
def test_cavity_bitwise(kind, p):
    steps = 300
    sc = _cavity(_medium(kind, p))
    sc.probes = [S.Probe(S.EZ, 4, 3, 1)]
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    for s in (g, o):
        s.set_field(S.EZ, _mode(s))
        s.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    assert np.array_equal(g.get_probes(), o.get_probes())


def test_lorentz_cavity_fp32_close():
    steps = 300
    sc = _cavity(_medium("lorentz", LORENTZ[0]))
    sc.probes = [S.Probe(S.EZ, 4, 3, 1)]
    g, o = _sim(sc, steps, 32), OracleSim(sc, steps)
    for s in (g, o):
        s.set_field(S.EZ, _mode(s))
        s.step(steps)
    a, b = g.get_probes()[0], o.get_probes()[0]
    print(f"Lorentz cavity FP32 probe error {np.abs(a - b).max() / np.abs(b).max():.3e}", flush=True)
    assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


def test_checkpoint_carries_lorentz_state():
    steps = 120
    sc = _cavity(_medium("lorentz", LORENTZ[1]))
    g = _sim(sc, steps)
    g.set_field(S.EZ, _mode(g))
    g.step(40)
    g.checkpoint(0)
    g.step(60)
    first = [g.get_field(c) for c in range(6)]
    g.restore(0)
    g.step(60)
    for c in range(6):
        assert np.array_equal(g.get_field(c), first[c]), c


def test_gradient_refuses_debye_and_lorentz():
    from adjoint_scenes import design_scene
    for med in (S.DebyMedium("d", 4.0, 2.0, 1e-11), S.LorentzMedium("l", 3.0, 1.5, 1e11, 1e10)):
        sc, _ = design_scene()
        dl = sc.cell_size
        corner = -13 * dl + 7 * dl  # cell (7, 7, 7): outside the CPML and the design box
        sc.objects = list(sc.objects) + [S.Object("m", S.Cube((corner,) * 3, (2 * dl,) * 3), med)]
        with pytest.raises(RuntimeError, match="Drude media only"):
            _sim(sc, 40).grad(checkpoint_every=10)


def test_cpp_lorentz_sphere_monostatic_against_mie(tmp_path):
    """The reference's Lorentz-sphere monostatic test
    (test/test_3d_fdtd.cpp:158-211) through the C++ API: a 15 nm sphere of
    LorentzMedium(2.25, 1, 4e16, 0.28e16) at dl = 0.5 nm, 1000 steps, FP64.
    sigma_b(f) = 4 pi (P_theta + P_phi) / (|E_inc(f)|^2 / 2 eta) against the
    Mie series of a sphere of eps(w) = 1 + 1.25 wp^2 / (wp^2 - w^2 - 2i w nu)
    over 2 .. 15 PHz: the resonance peak at the Mie frequency (6.8 PHz,
    within 3 %), the band's total within 1.5 dB, and the bins where Mie is
    within 10 dB of its peak within 2.5 dB on average (measured 2.0; single
    bins at the series' sharp dips miss by up to 5 dB, and above 11 PHz,
    where the sphere scatters 20 dB below its peak, the FDTD curve departs
    from the series by up to 20 dB)."""
    import os
    import subprocess
    from test_broadband_gpu import _gaussian_spectrum
    from test_dispersive_cpu import lorentz_eps, mie_sphere_backscatter
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "xfdtd_amd", "examples", "lorentz_sphere_monostatic_rcs")
    assert os.path.exists(exe), "build the library first (make -C xfdtd_amd/csrc)"
    out = tmp_path / "lorentz"
    steps = 1000
    print(f"running {exe} {out} {steps} 64", flush=True)
    r = subprocess.run([exe, str(out), str(steps), "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    a, dl = 15e-9, 15e-9 / 30
    pt, pp = np.load(out / "power_theta.npy"), np.load(out / "power_phi.npy")
    Ns = 2 * steps - 1
    dt = S.courant_dt(dl, 0.99)
    f = np.linspace(-1 / dt / 2, 1 / dt / 2, Ns)
    m = np.arange(Ns) - (steps - 1)
    band = (f > 2e15) & (f < 1.5e16)
    tau = 20 * dl / 6e8
    einc = _gaussian_spectrum(steps, dt, tau, 4.5 * tau, m[band])
    sigma = 4 * math.pi * (pt[band] + pp[band]) / (np.abs(einc) ** 2 / (2 * S.ETA_0))
    w = 2 * math.pi * f[band]
    k = w / S.C_0
    mie = np.array([mie_sphere_backscatter(kk * a, np.sqrt(lorentz_eps(ww, 2.25, 1.0, 4e16, 0.28e16)))
                    for kk, ww in zip(k, w)]) * math.pi * a * a
    err = 10 * np.log10(sigma / mie)
    bias = 10 * math.log10(sigma.sum() / mie.sum())
    peak_f, peak_m = f[band][np.argmax(sigma)], f[band][np.argmax(mie)]
    print(f"Lorentz sphere vs Mie over {band.sum()} bins: band total {bias:+.2f} dB, mean {np.abs(err).mean():.2f} dB, "
          f"worst {np.abs(err).max():.2f} dB; peak {peak_f / 1e15:.2f} PHz (Mie {peak_m / 1e15:.2f})", flush=True)
    strong = mie >= 0.1 * mie.max()
    print(f"  bins within 10 dB of the Mie peak: {strong.sum()}, mean {np.abs(err[strong]).mean():.2f} dB", flush=True)
    assert abs(bias) < 1.5, err
    assert strong.sum() >= 10 and np.abs(err[strong]).mean() < 2.5, err
    assert abs(peak_f - peak_m) < 0.03 * peak_m


def test_cpp_drude_sphere_monostatic_against_mie(tmp_path):
    """The reference's Drude-sphere monostatic test
    (test/test_3d_fdtd.cpp:213-270) through the same C++ program: a 3.75 mm
    sphere of DrudeMedium(1, 1.8e11, 2e10) at dl = radius / 30, FP64, against
    the Mie series of eps(w) = 1 - wp^2 / (w^2 + i w nu) over 5 .. 60 GHz.
    The reference runs 1000 steps, which leaves 26 bins in the band (2.1 GHz
    apart); 3000 steps give 0.7 GHz bins, so the comparison resolves the
    series' peak: the band's total within 1.5 dB, at least 10 bins within
    10 dB of the series' peak agreeing within 2.5 dB on average, and the
    backscatter peak at the series' peak frequency within 3 % or one bin."""
    import os
    import subprocess
    from test_broadband_gpu import _gaussian_spectrum
    from test_dispersive_cpu import mie_sphere_backscatter
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "xfdtd_amd", "examples", "lorentz_sphere_monostatic_rcs")
    out = tmp_path / "drude"
    steps = 3000
    print(f"running {exe} {out} {steps} 64 drude", flush=True)
    r = subprocess.run([exe, str(out), str(steps), "64", "drude"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    a = 3.75e-3
    dl = a / 30
    pt, pp = np.load(out / "power_theta.npy"), np.load(out / "power_phi.npy")
    Ns = 2 * steps - 1
    dt = S.courant_dt(dl, 0.99)
    f = np.linspace(-1 / dt / 2, 1 / dt / 2, Ns)
    m = np.arange(Ns) - (steps - 1)
    band = (f > 5e9) & (f < 6e10)
    tau = 20 * dl / 6e8
    einc = _gaussian_spectrum(steps, dt, tau, 4.5 * tau, m[band])
    sigma = 4 * math.pi * (pt[band] + pp[band]) / (np.abs(einc) ** 2 / (2 * S.ETA_0))
    w = 2 * math.pi * f[band]
    eps = 1 - 1.8e11 ** 2 / (w ** 2 + 1j * w * 2e10)
    mie = np.array([mie_sphere_backscatter(ww / S.C_0 * a, np.sqrt(e)) for ww, e in zip(w, eps)]) * math.pi * a * a
    err = 10 * np.log10(sigma / mie)
    bias = 10 * math.log10(sigma.sum() / mie.sum())
    strong = mie >= 0.1 * mie.max()
    peak_f, peak_m = f[band][np.argmax(sigma)], f[band][np.argmax(mie)]
    print(f"Drude sphere vs Mie over {band.sum()} bins: band total {bias:+.2f} dB, mean {np.abs(err).mean():.2f} dB, "
          f"worst {np.abs(err).max():.2f} dB; {strong.sum()} bins within 10 dB of the peak, mean "
          f"{np.abs(err[strong]).mean():.2f} dB; peak {peak_f / 1e9:.2f} GHz (Mie {peak_m / 1e9:.2f})", flush=True)
    assert abs(bias) < 1.5, err
    assert strong.sum() >= 10 and np.abs(err[strong]).mean() < 2.5, err
    assert abs(peak_f - peak_m) <= max(0.03 * peak_m, 1.01 * (f[1] - f[0]))


def debye_eps(w, eps_s, eps_inf, tau, sigma):
    """The Debye permittivity with conductivity in the exp(-i w t) convention
    (Im eps > 0): eps_inf + (eps_s - eps_inf) / (1 - i w tau) + i sigma / (w eps_0)."""
    return eps_inf + (eps_s - eps_inf) / (1 - 1j * w * tau) + 1j * sigma / (w * S.EPSILON_0)


def _debye_run(tmp_path, steps, radii):
    """The Debye-sphere program's backscatter cross-section and the Mie
    series over the bins of 50 .. 350 MHz (ka 0.26 .. 1.8, the first lobe of
    the backscatter)."""
    import os
    import subprocess
    from test_broadband_gpu import _gaussian_spectrum
    from test_dispersive_cpu import mie_sphere_backscatter
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "xfdtd_amd", "examples", "lorentz_sphere_monostatic_rcs")
    out = tmp_path / f"debye{radii}"
    print(f"running {exe} {out} {steps} 64 debye {radii}", flush=True)
    r = subprocess.run([exe, str(out), str(steps), "64", "debye", str(radii)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    a = 0.25
    dl = a / 30
    pt, pp = np.load(out / "power_theta.npy"), np.load(out / "power_phi.npy")
    assert np.isfinite(pt).all() and np.isfinite(pp).all()
    Ns = 2 * steps - 1
    dt = S.courant_dt(dl, 0.99)
    f = np.linspace(-1 / dt / 2, 1 / dt / 2, Ns)
    m = np.arange(Ns) - (steps - 1)
    band = (f > 5e7) & (f < 3.5e8)
    tau = 20 * dl / 6e8
    einc = _gaussian_spectrum(steps, dt, tau, 4.5 * tau, m[band])
    sigma = 4 * math.pi * (pt[band] + pp[band]) / (np.abs(einc) ** 2 / (2 * S.ETA_0))
    w = 2 * math.pi * f[band]
    eps = debye_eps(w, 1.16, 1.01, 6.497e-10, 2.95e-4)
    mie = np.array([mie_sphere_backscatter(ww / S.C_0 * a, np.sqrt(e)) for ww, e in zip(w, eps)]) * math.pi * a * a
    return f[band], sigma, mie


def test_cpp_debye_sphere_monostatic_against_mie(tmp_path):
    """The reference's Debye-sphere monostatic test (test/test_3d_fdtd.cpp:
    268-321) through the C++ program: a 0.25 m sphere of
    DebyMedium(1.16, 1.01, 6.497e-10, 2.95e-4) at dl = radius / 30, FP64,
    against the Mie series of eps(w) = 1.01 + 0.15 / (1 - i w tau) +
    i sigma / (w eps_0) over the first lobe of the backscatter (50 .. 350
    MHz, ka 0.26 .. 1.8; above it this weak scatterer's backscatter falls
    to the level of the staircased surface's own scattering).

    - In the reference's geometry (a 2.5-radius air cube, TFSF at index 22)
      the TFSF surface runs through the sphere, so the scattered field holds
      the incident wave's mismatch there; at the reference's 1000 steps the
      curve runs below Mie (measured over the 10 bins: band total -2.2 dB,
      bins -5.7 .. +1.6 dB).  Checked: the run, its peak inside the lobe,
      the band total within 5 dB.
    - With the sphere inside the total-field region (a 3.2-radius cube,
      4000 steps so the fields have decayed; 38 bins): every bin within
      1.5 dB (measured: worst 1.35, at the lowest bins), the band total
      within 0.5 dB (measured -0.13) and the peak at the series' peak
      frequency within 3 % or one bin (236 against 228 MHz)."""
    f, sigma, mie = _debye_run(tmp_path, 1000, 2.5)
    err = 10 * np.log10(sigma / mie)
    bias = 10 * math.log10(sigma.sum() / mie.sum())
    print(f"Debye sphere, reference geometry, vs Mie over {f.size} bins: band total {bias:+.2f} dB, "
          f"bins {err.min():+.2f} .. {err.max():+.2f} dB", flush=True)
    assert abs(bias) < 5, err
    assert f[0] < f[np.argmax(sigma)] < f[-1]
    f, sigma, mie = _debye_run(tmp_path, 4000, 3.2)
    err = 10 * np.log10(sigma / mie)
    bias = 10 * math.log10(sigma.sum() / mie.sum())
    peak_f, peak_m = f[np.argmax(sigma)], f[np.argmax(mie)]
    print(f"Debye sphere inside the TF region vs Mie over {f.size} bins: band total {bias:+.2f} dB, "
          f"worst {np.abs(err).max():.2f} dB; peak {peak_f / 1e6:.1f} MHz (Mie {peak_m / 1e6:.1f})", flush=True)
    assert f.size >= 30
    assert np.abs(err).max() < 1.5, err
    assert abs(bias) < 0.5, err
    assert abs(peak_f - peak_m) <= max(0.03 * peak_m, 1.01 * (f[1] - f[0]))
