"""The fused E+H step (xf_engine_set_sweep mode 1, k_sweep) against the
oracle and against the split E/H kernels.

- FP64: fields, NF-FF accumulators and probe samples bit-identical to the
  oracle for every fused launch geometry (1, 2 or 4 j rows per thread; 1 to
  32 x planes per workgroup), on a PEC sphere with TFSF and
  NF-FF at oblique incidence and on seeded random fields with kappa > 1 CPML
  on every face (every wall and slab range, no source).
- Fused tiles come in three kinds (plain, z slab and faces only, any):
  a 144^3 scene that has all three is bit-identical to the oracle in FP64,
  and the 400^3 dielectric-sphere run from seeded random fields is
  bit-identical to the split kernels in FP64.
- FP32: fused and split runs bit-identical.
- The double-buffered state: a checkpoint taken after an odd number of fused
  steps restores bit for bit; switching between the fused and split schemes
  mid-run continues exactly; engines with a current source fall back to the
  split kernels.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(sc, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(sc, steps, precision=precision)


def pec_oblique(n=36, pml=5):
    dl = 5e-3
    objs = [S.Object("pec", S.Sphere((0.002, -0.003, 0.001), 6 * dl), S.Material.pec()),
            S.Object("d", S.Cube((-0.03, 0.0, -0.02), (0.02, 0.02, 0.03)), S.Material("d", 3.0, 1.5, 0.02, 0.0))]
    sc = S.cube_domain_scene(n, dl, pml, cfl=0.9, objects=objs, tfsf_gap=pml + 4, nffft_gap=pml + 2,
                             freqs=[S.C_0 / (30 * dl), S.C_0 / (15 * dl)], theta=np.linspace(0, math.pi, 7),
                             phi=[0.0, 1.0], incidence=(0.7, 0.4, 0.3))
    sc.probes = [S.Probe(S.EX, 18, 18, 18), S.Probe(S.HZ, 9, 20, 30), S.Probe(S.EZ, 2, 3, 4)]
    return sc


def random_cpml():
    dl = 3e-3
    sc = S.Scene(cell_size=dl, cfl=0.95)
    sc.add_object(S.Object("box", S.Cube((-0.03, -0.027, -0.024), (0.06, 0.051, 0.045)), S.Material.air()))
    sc.add_object(S.Object("c", S.Cube((-0.01, -0.01, -0.01), (0.02, 0.015, 0.012)),
                           S.Material("m", 3.0, 2.0, 0.3, 1e3)))
    sc.boundaries = [S.PML(o, t, 3, 0.8, 0.0, 0.05, 3.0) for o, t in zip(range(6), (4, 5, 3, 4, 5, 3))]
    return sc


def _state(s):
    st = {f"f{c}": s.get_field(c) for c in range(6)}
    if s.scene.nffft is not None:
        st["nf"] = s.get_nffft()
    if s.scene.probes:
        st["probes"] = s.get_probes()
    return st


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, float(np.abs(a[k] - b[k]).max()))


GEOMETRIES = [(1, 0), (2, 0), (4, 0), (1, 1), (2, 3), (4, 5), (4, 16), (2, 32)]


@pytest.fixture(scope="module")
def pec_oracle(oracle_threads):
    sc = pec_oblique()
    steps = 80
    o = OracleSim(sc, steps)
    o.step(steps)
    return sc, steps, _state(o)


@pytest.mark.parametrize("rows,planes", GEOMETRIES)
def test_fused_tfsf_nffft_probes_bitwise(rows, planes, pec_oracle):
    sc, steps, ref = pec_oracle
    print(f"fused step, rows {rows}, planes {planes}: oblique PEC sphere 36^3 FP64", flush=True)
    g = _sim(sc, steps)
    g.set_sweep(1, rows, planes)
    assert g.sweep_active
    g.step(steps)
    _same(_state(g), ref)


@pytest.mark.parametrize("rows,planes", [(1, 0), (2, 7), (4, 0), (4, 2)])
def test_fused_random_fields_cpml_all_faces_bitwise(rows, planes, oracle_threads):
    sc = random_cpml()
    steps = 30
    o = OracleSim(sc, steps)
    g = _sim(sc, steps)
    g.set_sweep(1, rows, planes)
    rng = np.random.default_rng(4321)
    for c in range(6):
        f = rng.standard_normal(o.field_shape(c))
        o.set_field(c, f)
        g.set_field(c, f)
    print(f"fused step, rows {rows}, planes {planes}: random fields, CPML kappa 3 on all faces", flush=True)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c


@pytest.mark.parametrize("rows,planes", [(0, 0), (4, 8), (1, 32)])
def test_fused_all_tile_kinds_against_oracle(rows, planes, oracle_threads):
    # 144^3 with 8-cell CPML: the TFSF faces sit at 12 and 132, so the k chunk
    # [64, 128) and the x and y ranges between the faces hold plain tiles
    from xfdtd_amd.workloads import dielectric_sphere_scene
    sc = dielectric_sphere_scene(144, pml=8)
    steps = 40
    o = OracleSim(sc, steps)
    g = _sim(sc, steps)
    g.set_sweep(1, rows, planes)
    assert g.sweep_active
    rng = np.random.default_rng(144)
    for c in range(6):
        f = rng.standard_normal(o.field_shape(c))
        o.set_field(c, f)
        g.set_field(c, f)
    print(f"fused step, rows {rows}, planes {planes}: 144^3 dielectric sphere from random fields, FP64", flush=True)
    g.step(steps)
    o.step(steps)
    _same(_state(g), _state(o))


def test_fused_400_fp64_equals_split():
    from xfdtd_amd.workloads import dielectric_sphere_scene
    sc = dielectric_sphere_scene(400)
    steps = 12
    rng = np.random.default_rng(400)
    runs = []
    fields = None
    for mode in (1, 0):
        g = _sim(sc, steps)
        g.set_sweep(mode)
        assert g.sweep_active == (mode == 1)
        if fields is None:
            fields = [rng.standard_normal(g.field_shape(c)) for c in range(6)]
        for c in range(6):
            g.set_field(c, fields[c])
        print(f"400^3 dielectric sphere, FP64, {'fused' if mode else 'split'} scheme, {steps} steps", flush=True)
        g.step(steps)
        runs.append(_state(g))
        g.close()
    _same(runs[0], runs[1])


def test_fused_fp32_equals_split():
    sc = pec_oblique()
    steps = 90
    runs = []
    for mode in (1, 0):
        g = _sim(sc, steps, 32)
        g.set_sweep(mode)
        assert g.sweep_active == (mode == 1)
        g.step(steps)
        runs.append(_state(g))
    _same(runs[0], runs[1])


def test_checkpoint_after_odd_fused_steps_and_mode_switch(oracle_threads):
    sc = pec_oblique()
    T = 80
    o = OracleSim(sc, T)
    o.step(T)
    ref = _state(o)
    g = _sim(sc, T)
    g.set_sweep(1)
    g.step(33)                 # odd: the current state sits in the second copies
    g.checkpoint(0)
    g.step(10)
    g.restore(0)
    assert g.current_step == 33
    g.set_sweep(0)             # split kernels from the fused state
    g.step(12)
    g.set_sweep(1, 2, 5)       # and back
    g.step(T - 45)
    _same(_state(g), ref)


def test_current_source_uses_split_kernels():
    dl = 4e-3
    sc = S.cube_domain_scene(24, dl, 4, cfl=0.99)
    half = 12 * dl
    sc.sources.append(S.CurrentSource(S.Cube((-half + 10 * dl, -half + 12 * dl, -half + 12 * dl), (dl, dl, dl)),
                                      S.ZP, math.inf, S.GaussianWaveform(1e-10, 4e-10, 1.0)))
    g = _sim(sc, 10)
    assert not g.sweep_active  # the split kernels are the default
    g.set_sweep(1)
    assert not g.sweep_active
