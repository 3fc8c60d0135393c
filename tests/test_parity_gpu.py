"""GPU parity: the HIP engine (through the C-ABI) against the FP64 CPU oracle.

FP64 runs must agree to 1e-10 of each component's peak (they are expected to
be bit-identical: same operand order, no FMA contraction on either side);
NF-FF accumulators likewise; far-field quantities, whose face sums are
reduced in a different order on the GPU, to 1e-10 relative to their peak.
Probe samples must land on identical step indices.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10


def _gpu(scene, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(scene, steps, precision=precision)


def _rel(a, b):
    scale = np.abs(b).max()
    if scale == 0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)


def _compare_fields(g, o, tol=FIELD_TOL):
    errs = []
    for c in range(6):
        a, b = g.get_field(c), o.get_field(c)
        assert a.shape == b.shape
        errs.append(_rel(a, b))
    assert max(errs) <= tol, f"field rel errors {errs}"
    return errs


def pec_scene(n=40, pml=6, gap=4, radius_cells=8, incidence=(0.0, 0.0, 0.0), nf=True):
    dl = 5e-3
    objs = [S.Object("pec", S.Sphere((0, 0, 0), radius_cells * dl), S.Material.pec())]
    return S.cube_domain_scene(
        n, dl, pml, cfl=0.9, objects=objs, tfsf_gap=pml + gap,
        nffft_gap=(pml + gap - 2) if nf else None,
        freqs=[S.C_0 / (40 * dl), S.C_0 / (20 * dl)],
        theta=np.linspace(0, math.pi, 19), phi=[0.0, math.pi / 2], incidence=incidence)


def test_pec_sphere_fp64_bitwise(oracle_threads):
    sc = pec_scene()
    steps = 120
    g = _gpu(sc, steps)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    errs = _compare_fields(g, o)
    assert max(errs) == 0.0, f"FP64 path is expected to be bit-identical, got {errs}"
    ag, ao = g.get_nffft(), o.get_nffft()
    assert ag.shape == ao.shape
    assert np.array_equal(ag, ao)


@pytest.mark.parametrize("nf_gap", [-2, -1, 0])
def test_nffft_surface_against_tfsf_planes(nf_gap, oracle_threads):
    """The NF-FF z faces are sampled from records the H kernel leaves unless
    the TFSF face kernel corrects H on a record plane: an NF-FF box two cells
    outside the TFSF box takes the records, one cell outside (the TFSF H
    correction plane) and on the TFSF box itself take the fields.  All three
    give the oracle's accumulators bit for bit."""
    dl, pml, gap = 5e-3, 5, 5
    objs = [S.Object("pec", S.Sphere((0, 0, 0), 5 * dl), S.Material.pec())]
    sc = S.cube_domain_scene(34, dl, pml, cfl=0.9, objects=objs, tfsf_gap=pml + gap,
                             nffft_gap=pml + gap + nf_gap, freqs=[S.C_0 / (30 * dl)],
                             theta=np.linspace(0, math.pi, 7), phi=[0.0], incidence=(0.4, 0.2, 0.0))
    steps = 60
    g = _gpu(sc, steps)
    g.set_tiling(64, 4, 8, 0)  # the records need whole 64-lane k runs
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    assert max(_compare_fields(g, o)) == 0.0
    assert np.array_equal(g.get_nffft(), o.get_nffft())


def test_oblique_dielectric_fp64(oracle_threads):
    dl = 4e-3
    objs = [S.Object("diel", S.Sphere((0.004, -0.002, 0.0), 7 * dl), S.Material("d", 4.0, 1.0, 0.0, 0.0)),
            S.Object("lossy", S.Cube((-0.02, -0.02, -0.02), (0.012, 0.016, 0.008)),
                     S.Material("l", 2.5, 1.5, 0.05, 0.0))]
    sc = S.cube_domain_scene(36, dl, 5, cfl=0.99, objects=objs, tfsf_gap=9, nffft_gap=7,
                             freqs=[S.C_0 / (30 * dl)], theta=np.linspace(0, math.pi, 13),
                             phi=[0.3, 1.9], incidence=(0.6, 0.9, 0.3))
    steps = 90
    g = _gpu(sc, steps)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    errs = _compare_fields(g, o)
    assert max(errs) == 0.0, errs
    assert np.array_equal(g.get_nffft(), o.get_nffft())


def test_cpml_random_initial_fields_fp64(oracle_threads):
    """Seeded random fields everywhere (including the outer walls), CPML on
    all faces with kappa > 1, no source: exercises every Yee and CPML index
    range, the wall terms the CPML writes, and the kappa division."""
    dl = 3e-3
    sc = S.Scene(cell_size=dl, cfl=0.95)
    sc.add_object(S.Object("box", S.Cube((-0.03, -0.027, -0.024), (0.06, 0.051, 0.045)), S.Material.air()))
    sc.add_object(S.Object("c", S.Cube((-0.01, -0.01, -0.01), (0.02, 0.015, 0.012)),
                           S.Material("m", 3.0, 2.0, 0.3, 1e3)))
    sc.boundaries = [S.PML(o, t, 3, 0.8, 0.0, 0.05, 3.0) for o, t in zip(range(6), (4, 5, 3, 4, 5, 3))]
    steps = 40
    o = OracleSim(sc, steps)
    g = _gpu(sc, steps)
    rng = np.random.default_rng(1234)
    for c in range(6):
        f = rng.standard_normal(o.field_shape(c))
        o.set_field(c, f)
        g.set_field(c, f)
    g.step(steps)
    o.step(steps)
    errs = _compare_fields(g, o)
    assert max(errs) == 0.0, errs


def test_current_source_and_probes_fp64(oracle_threads):
    dl = 4e-3
    n, d = 40, 6
    sc = S.cube_domain_scene(n, dl, d, cfl=0.99, objects=[
        S.Object("d", S.Cube((-0.02, -0.02, -0.02), (0.04, 0.04, 0.04)), S.Material("d", 2.0, 1.0, 0.0, 0.0))])
    dt = S.courant_dt(dl, 0.99)
    tau = 20 * dt
    # one Ez cell at (12, 20, 20) in grid indices
    half = (n - 2 * d) * dl / 2 + d * dl
    x0 = -half + 12 * dl
    y0 = -half + 20 * dl
    z0 = -half + 20 * dl
    sc.sources.append(S.CurrentSource(S.Cube((x0, y0, z0), (dl, dl, dl)), S.ZP, math.inf,
                                      S.GaussianWaveform(tau, 4.5 * tau, 1.0)))
    sc.probes = [S.Probe(S.EZ, 12, 20, 20), S.Probe(S.EZ, 28, 20, 20), S.Probe(S.HY, 20, 20, 20)]
    steps = 150
    g = _gpu(sc, steps)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    errs = _compare_fields(g, o)
    assert max(errs) == 0.0, errs
    pg, po = g.get_probes(), o.get_probes()
    assert np.array_equal(pg, po)
    assert np.abs(po[1]).max() > 0


def test_farfield_fp64(oracle_threads):
    sc = pec_scene(n=36, pml=5, gap=4, radius_cells=6)
    steps = 160
    g = _gpu(sc, steps)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    fg, fo = g.farfield(), o.farfield()
    for k in fo:
        err = _rel(fg[k], fo[k])
        assert err <= 1e-10, (k, err)


def test_fp32_error_vs_fp64_oracle(oracle_threads):
    sc = pec_scene(n=40, pml=6, gap=4, radius_cells=8, incidence=(0.4, 0.2, 0.0))
    steps = 150
    g = _gpu(sc, steps, precision=32)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    errs = [_rel(g.get_field(c), o.get_field(c)) for c in range(6)]
    # FP32 fields against the FP64 oracle: error grows like sqrt(steps) ulps
    ag, ao = g.get_nffft(), o.get_nffft()
    print(f"FP32 field errors {['%.2e' % e for e in errs]}, NF-FF {_rel(ag, ao):.2e}", flush=True)
    assert max(errs) < 1e-5, errs
    assert _rel(ag, ao) < 1e-5


def test_step_beyond_horizon_is_refused():
    sc = pec_scene(n=28, pml=4, gap=3, radius_cells=4, nf=False)
    g = _gpu(sc, 5)
    g.step(5)
    with pytest.raises(RuntimeError):
        g.step(1)


@pytest.mark.parametrize("tiling", [(64, 4, 16, 0), (256, 0, 16, 0), (64, 0, 3, 2), (128, 2, 5, 0),
                                    (64, 1, 0, 0), (256, 0, 1, 6), (64, 4, 2, 4), (64, 4, 8, 8),
                                    (128, 2, 3, 8), (64, 4, 8, 16), (64, 2, 5, 18)])
def test_launch_tilings_bitwise(tiling, oracle_threads):
    """Every launch geometry of the split curl kernels (2-D tiles, flat plane
    runs, global or LDS tables, whole-axis marches, every tile order) gives
    the same bits."""
    sc = pec_scene(n=34, pml=5, gap=4, radius_cells=6, incidence=(0.5, 0.3, 0.1))
    steps = 70
    g = _gpu(sc, steps)
    g.set_sweep(0)
    g.set_tiling(*tiling)
    o = OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    errs = _compare_fields(g, o)
    assert max(errs) == 0.0, errs
    assert np.array_equal(g.get_nffft(), o.get_nffft())
