"""Pins of the oracle's adjoint (the checker of the GPU adjoint).

- Transpose identity: for the field map L of n steps (Yee + CPML with
  kappa > 1 on every face, random initial fields everywhere, walls
  included), <L x, y> = <x, L^T y> to rounding.  A wrong index, sign or
  coefficient in any transposed update breaks it.
- Central finite differences of J = sum |DFT Ez|^2 over a plane with respect
  to the permittivity of single design cells agree with the adjoint
  gradient.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S


def box_scene(n=(14, 12, 13), dl=3e-3, pml=3, kappa=3.0):
    sc = S.Scene(cell_size=dl, cfl=0.95)
    size = tuple(v * dl for v in n)
    sc.add_object(S.Object("box", S.Cube(tuple(-s / 2 for s in size), size), S.Material.air()))
    sc.add_object(S.Object("c", S.Cube((-0.006, -0.005, -0.004), (0.012, 0.009, 0.01)),
                           S.Material("m", 3.0, 2.0, 0.2, 0.0)))
    sc.boundaries = [S.PML(o, pml, 3, 0.8, 0.0, 0.05, kappa) for o in range(6)]
    return sc


def test_adjoint_transpose_identity(oracle_threads):
    steps = 25
    sc = box_scene()
    o = OracleSim(sc, steps)
    rng = np.random.default_rng(7)
    x0 = [rng.standard_normal(o.field_shape(c)) for c in range(6)]
    y = [rng.standard_normal(o.field_shape(c)) for c in range(6)]
    for c in range(6):
        o.set_field(c, x0[c])
    o.step(steps)
    lhs = sum(float(np.sum(o.get_field(c) * y[c])) for c in range(6))
    a = OracleSim(sc, steps)
    ly = a.adjoint_run(steps, y)
    rhs = sum(float(np.sum(x0[c] * ly[c])) for c in range(6))
    scale = sum(float(np.sum(np.abs(x0[c] * ly[c]))) for c in range(6))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


def test_adjoint_transpose_identity_fails_for_wrong_adjoint(oracle_threads):
    """The identity is sharp: one extra reverse step (a plausible off-by-one
    in the reverse loop) breaks it by many orders of magnitude."""
    steps = 12
    sc = box_scene()
    o = OracleSim(sc, steps + 1)
    rng = np.random.default_rng(8)
    x0 = [rng.standard_normal(o.field_shape(c)) for c in range(6)]
    y = [rng.standard_normal(o.field_shape(c)) for c in range(6)]
    for c in range(6):
        o.set_field(c, x0[c])
    o.step(steps)
    lhs = sum(float(np.sum(o.get_field(c) * y[c])) for c in range(6))
    ly = OracleSim(sc, steps + 1).adjoint_run(steps + 1, y)
    rhs = sum(float(np.sum(x0[c] * ly[c])) for c in range(6))
    scale = sum(float(np.sum(np.abs(x0[c] * ly[c]))) for c in range(6))
    assert abs(lhs - rhs) > 1e-6 * scale


def design_scene(n=26, dl=4e-3, pml=5, steps=160, seed=0xC4, lo=11, hi=15):
    """A small version of the adjoint workload: soft z current source, design
    box of random eps_r in [1, 4], objective on an Ez plane past the box."""
    sc = S.cube_domain_scene(n, dl, pml, cfl=0.99)
    dt = S.courant_dt(dl, 0.99)
    tau = 20 * dt
    half = n * dl / 2
    x0, y0, z0 = -half + 8 * dl, -half + (n // 2) * dl, -half + (n // 2) * dl
    sc.sources.append(S.CurrentSource(S.Cube((x0, y0, z0), (dl, dl, dl)), S.ZP, math.inf,
                                      S.GaussianWaveform(tau, 4.5 * tau, 1.0)))
    rng = np.random.default_rng(seed)
    m = hi - lo
    sc.design = S.Design((lo, hi, lo, hi, lo, hi), rng.uniform(1.0, 4.0, m ** 3).astype(np.float32))
    sc.dft_plane = S.DftPlane(S.EZ, 0, 18, pml + 2, n - pml - 2, pml + 2, n - pml - 2, S.C_0 / (20 * dl))
    return sc, steps


def _J(sc, steps):
    o = OracleSim(sc, steps)
    o.step(steps)
    return float(np.sum(np.abs(o.get_dft()) ** 2))


def test_adjoint_gradient_matches_central_differences(oracle_threads):
    sc, steps = design_scene()
    J, g = OracleSim(sc, steps).grad()
    assert J > 0 and np.abs(g).max() > 0
    assert J == pytest.approx(_J(sc, steps), rel=1e-14)
    eps = np.asarray(sc.design.eps_r, dtype=np.float32).copy()
    m = sc.design.box[1] - sc.design.box[0]
    rng = np.random.default_rng(3)
    for flat in rng.choice(eps.size, 6, replace=False):
        e0 = float(eps[flat])
        hi, lo = np.float32(e0 + 1e-3), np.float32(e0 - 1e-3)
        vals = []
        for v in (hi, lo):
            e = eps.copy()
            e[flat] = v
            sc.design = S.Design(sc.design.box, e)
            vals.append(_J(sc, steps))
        sc.design = S.Design(sc.design.box, eps)
        fd = (vals[0] - vals[1]) / (float(hi) - float(lo))
        ga = g.reshape(-1)[flat]
        assert ga == pytest.approx(fd, rel=2e-5, abs=1e-6 * np.abs(g).max()), (np.unravel_index(flat, (m,) * 3), ga, fd)
