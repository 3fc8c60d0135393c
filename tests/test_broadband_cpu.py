"""Pins of the oracle's broadband NF-FF (nffft/nffft_broadband.cpp), the
monostatic radar cross-section path.

- The stored face-current series reproduce the NffftFd running-DFT
  accumulators of the same run (nffft_fd.cpp:91-203), which the bistatic
  Mie test pins: samples[n, c, q] weighted by dt e^{-i 2 pi f n dt} (magnetic
  currents) and dt e^{-i 2 pi f (n + 1/2) dt} (electric currents) and summed
  over n give every accumulator.  A wrong series layout, sign or step index
  fails this.
- The surface integrals equal NffftFd's at the bin frequencies
  f = m / (Ns dt), Ns = 2T - 1, once both use the same wave number: bin
  m = q - (T - 1) of fftshift(fft(series)) is exactly NffftFd's DFT without
  dt, the electric currents carry NffftFd's extra half-step phase
  e^{-i pi f dt}, and the broadband delays are measured from the box centre
  (grid/grid_box.cpp:101-105) where NffftFd's are absolute.  So
  N_fd = dt e^{-i pi f dt} e^{i k r.c dl} N_bb and L_fd = dt e^{i k r.c dl} L_bb,
  checked for positive and negative m.  A swapped current, a wrong target
  axis, a wrong delay offset, a conjugated bin or a misplaced fftshift fails.
- The bins' frequencies are numpy.linspace(-fs/2, fs/2, 2T-1), fs = 1/dt.
- The far-zone formulas (nffft_broadband.cpp:326-343) applied to NffftFd's
  projected integrals give NffftFd's far field (the Mie-pinned formulas of
  nffft_fd.cpp:379-421).
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim, bb_farfield
from xfdtd_amd import scene as S


def small_scene(nffft=None):
    """24^3 incl. 4-cell CPML, a dielectric sphere, oblique TFSF."""
    dl = 5e-3
    sph = S.Object("d", S.Sphere((1e-3, -2e-3, 0.5e-3), 4 * dl), S.Material("d", 3.0, 1.0, 0.0, 0.0))
    sc = S.cube_domain_scene(24, dl, 4, cfl=0.9, objects=[sph], tfsf_gap=6,
                             incidence=(0.7, 0.4, 0.3), tau_steps=6.0)
    sc.nffft = nffft
    return sc


THETA, PHI = 2.1, 0.9


def test_bb_frequencies_are_linspace():
    T = 40
    o = OracleSim(small_scene(S.NffftBroadBand(5, 5, 5, THETA, PHI)), T)
    f, k = o.bb_frequencies()
    fs = 1 / o.dt
    assert f.shape == (2 * T - 1,)
    assert np.allclose(f, np.linspace(-fs / 2, fs / 2, 2 * T - 1), rtol=0, atol=1e-15 * fs)
    assert np.array_equal(k, 2 * math.pi * f / S.C_0)


def test_bb_samples_reproduce_fd_accumulators(oracle_threads):
    T = 70
    freqs = [1.1e9, 4.0e9, -2.5e9]
    fd = S.NffftFd(5, 5, 5, freqs, [THETA], [PHI])
    o = OracleSim(small_scene(fd), T, broadband=S.NffftBroadBand(5, 5, 5, THETA, PHI))
    o.step(T)
    smp = o.get_bb_samples()                      # [T][8][Q]
    acc = o.get_nffft()
    dt = o.dt
    n = np.arange(T)
    f = np.asarray(freqs)[:, None]
    wm = dt * np.exp(-2j * np.pi * f * n * dt)            # [nf][T]
    wj = dt * np.exp(-2j * np.pi * f * (n + 0.5) * dt)
    nb = 14
    cells = [nb * nb] * 3
    off, qoff = 0, 0
    assert np.abs(smp).max() > 0
    for pair in range(3):
        for c in range(8):
            w = wm if c % 4 < 2 else wj
            expect = w @ smp[:, c, qoff:qoff + cells[pair]]   # [nf][cells]
            got = acc[off:off + len(freqs) * cells[pair]].reshape(len(freqs), cells[pair])
            off += len(freqs) * cells[pair]
            scale = np.abs(got).max()
            assert scale > 0
            assert np.abs(got - expect).max() <= 1e-12 * scale, (pair, c)
        qoff += cells[pair]
    assert off == acc.size and qoff == smp.shape[2]


def test_bb_sums_equal_fd_surface_integrals(oracle_threads):
    T = 90
    Ns = 2 * T - 1
    bb = S.NffftBroadBand(5, 5, 5, THETA, PHI)
    o0 = OracleSim(small_scene(bb), T)
    dt = o0.dt
    ms = [3, 11, 26, -7]
    fbin = [m / (Ns * dt) for m in ms]
    fd = S.NffftFd(5, 5, 5, fbin, [THETA], [PHI])
    o = OracleSim(small_scene(fd), T, broadband=bb)
    o.step(T)
    ff = o.farfield()
    # every bin at its own FFT frequency's wave number
    fall = (np.arange(Ns) - (T - 1)) / (Ns * dt)
    kall = 2 * math.pi * fall / S.C_0
    sums = o.bb_sums(k=kall)
    pr = bb_farfield(sums, (fall, kall), THETA, PHI)
    li, nb = 5, 14
    c = li + nb // 2
    rx, ry, rz = math.sin(THETA) * math.cos(PHI), math.sin(THETA) * math.sin(PHI), math.cos(THETA)
    for u, m in enumerate(ms):
        q = m + T - 1
        f, k = fbin[u], kall[q]
        shift = np.exp(1j * k * (rx * c + ry * c + rz * c) * o.scene.cell_size)
        for name, half in (("n_theta", True), ("n_phi", True), ("l_theta", False), ("l_phi", False)):
            got = ff[name][u, 0, 0]
            expect = dt * shift * pr[name][q] * (np.exp(-1j * np.pi * f * dt) if half else 1.0)
            assert abs(got) > 0
            assert abs(got - expect) <= 1e-10 * abs(got), (name, m, got, expect)


def test_bb_far_zone_formulas_equal_fd(oracle_threads):
    """bb_farfield's projections and far-zone formulas on NffftFd's
    integrals, handed in as Cartesian vectors N = N_theta theta^ + N_phi phi^
    (which project back exactly), give NffftFd's far field."""
    T = 60
    freqs = [2.0e9, 5.0e9]
    theta, phi = [0.4, 1.9], [0.3, 2.5]
    o = OracleSim(small_scene(S.NffftFd(5, 5, 5, freqs, theta, phi)), T)
    o.step(T)
    ff = o.farfield()
    k = 2 * math.pi * np.asarray(freqs) / S.C_0
    for t, th in enumerate(theta):
        for p, ph in enumerate(phi):
            th_hat = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
            ph_hat = np.array([-math.sin(ph), math.cos(ph), 0.0])
            nv = th_hat[:, None] * ff["n_theta"][:, t, p] + ph_hat[:, None] * ff["n_phi"][:, t, p]
            lv = th_hat[:, None] * ff["l_theta"][:, t, p] + ph_hat[:, None] * ff["l_phi"][:, t, p]
            r = bb_farfield(np.concatenate([nv, lv]), (np.asarray(freqs), k), th, ph)
            for name in ("e_theta", "e_phi", "h_theta", "h_phi", "power_theta", "power_phi"):
                ref = ff[name][:, t, p]
                assert np.abs(r[name] - ref).max() <= 1e-12 * np.abs(ref).max(), (name, t, p)


def test_bb_host_plan_and_bad_box():
    """The host plans a broadband surface without a device and refuses one
    that does not fit the grid (nffft/nffft.cpp:22-34)."""
    from xfdtd_amd.sim import Simulation
    Simulation(small_scene(S.NffftBroadBand(5, 5, 5, THETA, PHI)), 20, plan_only=True).close()
    with pytest.raises(RuntimeError, match="NF-FF box does not fit"):
        Simulation(small_scene(S.NffftBroadBand(12, 5, 5, THETA, PHI)), 20, plan_only=True)
