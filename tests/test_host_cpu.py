"""CPU checks of the product's C++ host layer and its C-ABI (no GPU).

* The library loads and exports every function include/*.h declares.
* The CPU plan (Simulation::plan) - material index per cell plus one-
  dimensional coefficient tables - reproduces the oracle's 18 per-cell
  coefficient arrays (fdtd_basic_coff.cpp:93-121 after the CPML kappa
  division and the CurrentSource correction) bit for bit.
* Invalid configurations fail with the reference's exceptions as messages.
"""
import math
import os
import re

import numpy as np
import pytest

from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in ["include/xfdtd_engine.h", "include/xfdtd_sim.h"]:
        txt = open(os.path.join(ROOT, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(xf_[a-z_0-9]+)\s*\(", txt):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import xfdtd_amd
    from xfdtd_amd._lib import ENGINE_SYMBOLS, SIM_SYMBOLS
    L = xfdtd_amd.lib()
    declared = _declared_functions()
    assert len(declared) >= 40
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(set(ENGINE_SYMBOLS + SIM_SYMBOLS)) == declared


def _scene_with_everything():
    dl = 3e-3
    sc = S.Scene(cell_size=dl, cfl=0.95)
    sc.add_object(S.Object("box", S.Cube((-0.03, -0.027, -0.024), (0.06, 0.051, 0.045)), S.Material.air()))
    sc.add_object(S.Object("c", S.Cube((-0.01, -0.01, -0.01), (0.02, 0.015, 0.012)),
                           S.Material("m", 3.0, 2.0, 0.3, 1e3)))
    sc.add_object(S.Object("s", S.Sphere((0.003, 0.0, -0.002), 0.008), S.Material("s", 5.5, 1.0, 0.0, 0.0)))
    sc.add_object(S.Object("pec", S.Sphere((-0.006, 0.004, 0.005), 0.004), S.Material.pec()))
    sc.boundaries = [S.PML(o, t, 3, 0.8, 0.0, 0.05, 3.0) for o, t in zip(range(6), (4, 5, 3, 4, 5, 3))]
    sc.sources.append(S.CurrentSource(S.Cube((0.0, 0.0, 0.0), (dl, dl, dl)), S.ZP, 50.0,
                                      S.GaussianWaveform(1e-10, 4e-10)))
    sc.sources.append(S.CurrentSource(S.Cube((-0.009, 0.003, 0.0), (dl, 2 * dl, dl)), S.XN, math.inf,
                                      S.GaussianWaveform(1e-10, 4e-10, 2.0)))
    return sc


def drude_scene(n=30, dl=4e-3, pml=5):
    """Two Drude objects (a sphere and a lossy cube) and a later plain cube
    that covers part of the sphere: those cells keep the sphere's Drude E
    update with the cube's standard coefficients."""
    wp = 2 * math.pi * 20e9
    objs = [S.Object("plasma", S.Sphere((0.002, -0.004, 0.0), 7 * dl), S.DrudeMedium("p", 1.0, wp, 2 * math.pi * 1e9)),
            S.Object("metal", S.Cube((-0.036, 0.01, -0.02), (0.02, 0.012, 0.016)),
                     S.DrudeMedium("q", 2.0, 0.6 * wp, 2 * math.pi * 3e9, 0.05, 0.0)),
            S.Object("cap", S.Cube((0.004, -0.03, -0.03), (0.03, 0.06, 0.06)), S.Material("c", 3.0, 1.2, 0.01, 0.0))]
    return S.cube_domain_scene(n, dl, pml, cfl=0.99, objects=objs, tfsf_gap=pml + 2, nffft_gap=pml + 1,
                               freqs=[S.C_0 / (20 * dl)], theta=[0.0, 1.0], phi=[0.0])


def lumped_scene(waveform=None, steps_hint=400):
    """The reference's basic lumped-element circuit
    (examples/basic_lumped_element.cpp): a 14 x 8 x 10 mm air box at 1 mm,
    two PEC plates (ObjectPlane) at z = 0 and 4 mm joined by a voltage
    source and a 150 ohm resistor, with a capacitor, an inductor, a current
    source, a later object covering part of the upper plate, and voltage and
    current monitors of the circuit."""
    size = 1e-3
    cube = lambda o, d: S.Cube(tuple(x * size for x in o), tuple(x * size for x in d))
    sc = S.Scene(cell_size=size, cfl=0.99)
    sc.add_object(S.Object("domain", cube((-3, -3, -3), (14, 8, 10)), S.Material.air()))
    sc.add_object(S.ObjectPlane("zn", cube((0, 0, 0), (8, 2, 0)), S.Material.pec()))
    sc.add_object(S.ObjectPlane("zp", cube((0, 0, 4), (8, 2, 0)), S.Material.pec()))
    sc.add_object(S.Object("patch", cube((6, 0, 4), (2, 2, 1)), S.Material("d", 3.0, 1.0, 0.01, 0.0)))
    w = waveform or S.SinusoidalWaveform(1.0, 5e10, 0.0)
    sc.sources = [S.VoltageSource(cube((0, 0, 0), (1, 2, 4)), S.ZP, 50.0, w),
                  S.Resistor(cube((7, 0, 0), (1, 2, 4)), S.AX_Z, 150.0),
                  S.Capacitor(cube((2, 0, 0), (1, 1, 4)), S.AX_Z, 2e-13),
                  # the inductor's explicit update is stable for dt^2 dc / (L' da db eps) <= 4
                  S.Inductor(cube((5, 1, 0), (1, 1, 4)), S.AX_Z, 3e-9),
                  S.CurrentSource(cube((3, -2, 1), (1, 1, 2)), S.YN, 1e5,
                                  S.CosineModulatedGaussianWaveform(1e-3, 1e-11, 4e-11, 6e10))]
    sc.monitors = [S.VoltageMonitor(cube((4, 0, 0), (1, 2, 4)), S.ZP, "", "circuit_voltage.npy"),
                   S.CurrentMonitor(cube((4, 0, 3), (1, 2, 1)), S.XP, "", "circuit_current.npy"),
                   S.CurrentMonitor(cube((2, 0, 1), (0, 1, 2)), S.ZN, "", "cap_current.npy"),
                   S.VoltageMonitor(cube((5, 1, 0), (1, 1, 4)), S.ZN, "", "ind_voltage.npy")]
    return sc


@pytest.mark.parametrize("builder", ["everything", "pec_example", "drude", "lumped", "cylinder_2d"])
def test_plan_reproduces_reference_coefficient_arrays(builder):
    from xfdtd_amd.sim import Simulation
    if builder == "cylinder_2d":
        # a dielectric then a PEC cylinder on a 2-D grid with CPML on x and y
        dl = 0.75e-3
        sc = S.cylinder_2d_scene(60, 50, 8, dl, 8.3 * dl, tfsf_gap=14, nffft_gap=11)
        sc.objects.insert(1, S.Object("d", S.Cylinder(S.AX_Z, (20 * dl, 20 * dl, 0.0), 9.6 * dl, 0.0),
                                      S.Material("d", 3.5, 1.0, 0.2, 0.0)))
    elif builder == "everything":
        sc = _scene_with_everything()
    elif builder == "lumped":
        sc = lumped_scene()
    elif builder == "drude":
        sc = drude_scene()
    else:
        dl = 5e-3
        sc = S.cube_domain_scene(40, dl, 8, cfl=0.9,
                                 objects=[S.Object("pec", S.Sphere((0, 0, 0), 10 * dl), S.Material.pec())],
                                 tfsf_gap=12, nffft_gap=10, freqs=[1e9], theta=[0.0], phi=[0.0])
    g = Simulation(sc, 20, plan_only=True)
    o = OracleSim(sc, 20)
    assert g.dims == o.dims
    assert g.dt == o.dt
    for w in range(18):
        a, b = g.plan_coeff(w), o.get_coeff(w)
        assert a.shape == b.shape
        assert np.array_equal(a, b), (w, np.abs(a - b).max())


def test_plan_design_box_coefficients():
    """Per-cell design permittivity: the plan's E arrays at design cells (and
    everything else) equal the oracle's per-cell arrays bit for bit."""
    from xfdtd_amd.sim import Simulation
    dl = 4e-3
    sc = S.cube_domain_scene(24, dl, 4, cfl=0.99, objects=[
        S.Object("m", S.Cube((-0.02, -0.02, -0.02), (0.03, 0.04, 0.04)), S.Material("m", 2.0, 1.5, 0.1, 0.0))])
    rng = np.random.default_rng(11)
    sc.design = S.Design((8, 13, 9, 12, 7, 15), rng.uniform(1, 4, 5 * 3 * 8).astype(np.float32))
    g = Simulation(sc, 10, plan_only=True)
    o = OracleSim(sc, 10)
    for w in range(18):
        a, b = g.plan_coeff(w), o.get_coeff(w)
        assert np.array_equal(a, b), (w, np.abs(a - b).max())


def test_invalid_configurations_raise():
    from xfdtd_amd.sim import Simulation
    dl = 5e-3
    # TFSF box inside the CPML slab
    sc = S.cube_domain_scene(30, dl, 8, cfl=0.9, tfsf_gap=6)
    with pytest.raises(RuntimeError, match="TFSF"):
        Simulation(sc, 10, plan_only=True)
    # two boundaries on one face
    sc = S.cube_domain_scene(30, dl, 4, cfl=0.9)
    sc.boundaries.append(S.PML(S.XN, 4))
    with pytest.raises(RuntimeError, match="same face"):
        Simulation(sc, 10, plan_only=True)
    # no objects
    sc = S.Scene(cell_size=dl)
    with pytest.raises(RuntimeError, match="no objects"):
        Simulation(sc, 10, plan_only=True)


def test_2d_refusals():
    from xfdtd_amd.sim import Simulation
    dl = 0.75e-3
    sc = S.cylinder_2d_scene(40, 40, 6, dl, 0, tfsf_gap=None, nffft_gap=None)
    sc.tfsf = S.TFSF3D(10, 10, 1, 1.0, 1.0, 0.5, 0.0, S.GaussianWaveform(1e-11, 4e-11))
    with pytest.raises(RuntimeError, match="TFSF2D"):
        Simulation(sc, 10, plan_only=True)
    sc = S.cylinder_2d_scene(40, 40, 6, dl, 0, tfsf_gap=12, nffft_gap=None)
    sc.nffft = S.NffftFd(8, 8, 0, [1e9], [1.0], [0.0])
    with pytest.raises(RuntimeError, match="2-D grid takes"):
        Simulation(sc, 10, plan_only=True)
    sc = S.cylinder_2d_scene(40, 40, 6, dl, 0, tfsf_gap=12, nffft_gap=8)
    sc.nffft.distance_z = 1
    with pytest.raises(RuntimeError, match="NF-FF box does not fit"):
        Simulation(sc, 10, plan_only=True)
    sc = S.cube_domain_scene(20, 5e-3, 4, tfsf_gap=None)
    sc.tfsf = S.TFSF2D(6, 6, 0.3, 1.0, S.GaussianWaveform(1e-11, 4e-11))
    with pytest.raises(RuntimeError, match="TFSF2D needs a 2-D grid"):
        Simulation(sc, 10, plan_only=True)
    Simulation(S.cylinder_2d_scene(40, 40, 6, dl, 5 * dl, tfsf_gap=12, nffft_gap=8), 10, plan_only=True).close()


def test_init_without_device_fails_loudly():
    """There is no CPU fallback: without a gfx950 device the engine refuses."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from xfdtd_amd.sim import Simulation
    sc = S.cube_domain_scene(20, 5e-3, 4, cfl=0.9)
    with pytest.raises(RuntimeError, match="HIP device|gfx950|hip"):
        Simulation(sc, 10)


def test_plan_refuses_unsupported_drude_placements():
    from xfdtd_amd.sim import Simulation
    dl = 4e-3
    wp = 2 * math.pi * 20e9
    # a Drude object reaching into the CPML
    sc = S.cube_domain_scene(24, dl, 4, cfl=0.99, objects=[
        S.Object("p", S.Cube((-0.05, -0.01, -0.01), (0.06, 0.02, 0.02)), S.DrudeMedium("p", 1.0, wp, 1e9))])
    with pytest.raises(RuntimeError, match="CPML"):
        Simulation(sc, 10, plan_only=True)
    # a current source on a Drude cell
    sc = S.cube_domain_scene(24, dl, 4, cfl=0.99, objects=[
        S.Object("p", S.Cube((-0.01, -0.01, -0.01), (0.02, 0.02, 0.02)), S.DrudeMedium("p", 1.0, wp, 1e9))])
    sc.sources.append(S.CurrentSource(S.Cube((0.0, 0.0, 0.0), (dl, dl, dl)), S.ZP, math.inf,
                                      S.GaussianWaveform(1e-10, 4e-10)))
    with pytest.raises(RuntimeError, match="Drude"):
        Simulation(sc, 10, plan_only=True)


def test_plan_refuses_bad_monitors(tmp_path):
    from xfdtd_amd.sim import Simulation
    dl = 5e-3
    base = lambda: S.cube_domain_scene(24, dl, 4, cfl=0.9)
    half = 12 * dl
    # an XY monitor two cells thick (field_monitor.cpp:34-37)
    sc = base()
    sc.monitors.append(S.TimeDomainFieldMonitor(S.Cube((-half, -half, 0.0), (2 * half, 2 * half, 2 * dl)),
                                                S.XY, S.EZ, str(tmp_path), "a.npy"))
    with pytest.raises(RuntimeError, match="invalid grid box"):
        Simulation(sc, 5, plan_only=True)
    # a YZ monitor of Ex at i = nx, outside Ex's (nx, ny+1, nz+1) array
    sc = base()
    sc.monitors.append(S.TimeDomainFieldMonitor(S.Cube((half, -half, -half), (0.0, dl, dl)),
                                                S.YZ, S.EX, str(tmp_path), "b.npy"))
    with pytest.raises(RuntimeError, match="outside the component"):
        Simulation(sc, 5, plan_only=True)
    # a movie monitor with plot_step 0
    sc = base()
    mon = S.TimeDomainFieldMonitor(S.Cube((-half, -half, 0.0), (dl, dl, 0.0)), S.XY, S.EZ, str(tmp_path))
    sc.monitors.append(S.MovieMonitor(mon, 5, 0))
    with pytest.raises(RuntimeError, match="plot_step"):
        Simulation(sc, 5, plan_only=True)
    # a good one plans
    sc = base()
    sc.monitors.append(S.MovieMonitor(mon, 5, 2))
    Simulation(sc, 5, plan_only=True)


def test_bad_lumped_and_monitor_placements():
    """The host refuses a current monitor whose loop leaves the grid, a plane
    object that is not a plane, a lumped element with no cells and an
    inductor inside a CPML slab."""
    from xfdtd_amd.sim import Simulation
    dl = 1e-3
    def base():
        sc = S.Scene(cell_size=dl, cfl=0.99)
        sc.add_object(S.Object("box", S.Cube((0, 0, 0), (10 * dl, 10 * dl, 10 * dl)), S.Material.air()))
        return sc
    sc = base()
    sc.monitors = [S.CurrentMonitor(S.Cube((0, 0, 2 * dl), (dl, dl, dl)), S.ZP)]
    with pytest.raises(RuntimeError, match="current monitor loop outside the grid"):
        Simulation(sc, 10, plan_only=True)
    sc = base()
    sc.add_object(S.ObjectPlane("p", S.Cube((dl, dl, dl), (dl, dl, dl)), S.Material.pec()))
    with pytest.raises(RuntimeError, match="is not a plane"):
        Simulation(sc, 10, plan_only=True)
    sc = base()
    sc.sources = [S.Resistor(S.Cube((dl, dl, dl), (dl, 0.0, dl)), S.AX_Z, 50.0)]
    with pytest.raises(RuntimeError, match="no cells"):
        Simulation(sc, 10, plan_only=True)
    sc = S.cube_domain_scene(20, dl, 4, cfl=0.99)
    half = 10 * dl
    sc.sources = [S.Inductor(S.Cube((-half + dl, 0, 0), (dl, dl, dl)), S.AX_Z, 1e-9)]
    with pytest.raises(RuntimeError, match="inductor inside a CPML slab"):
        Simulation(sc, 10, plan_only=True)


def test_check_run_writes_grid_and_object_descriptions(tmp_path, monkeypatch):
    """Simulation::checkRun (simulation.cpp:98-113): the CPU set-up, then the
    grid size in "simulation_check" and every object's description
    (object.cpp:43-47, cube.cpp:16-21, sphere.cpp:17-21, material.cpp:18-27,
    numbers in std::to_string's %f) in "simulation_object_check"."""
    from xfdtd_amd.sim import Simulation
    dl = 5e-3
    sc = S.Scene(cell_size=dl, cfl=0.9)
    sc.add_object(S.Object("domain", S.Cube((-0.05, -0.04, -0.03), (0.1, 0.08, 0.06)), S.Material.air()))
    sc.add_object(S.Object("ball", S.Sphere((0.01, 0.0, -0.005), 0.02), S.Material("glass", 2.25, 1.0, 0.01, 0.0)))
    s = Simulation(sc, 50, plan_only=True)
    monkeypatch.chdir(tmp_path)
    s.check_run()
    nx, ny, nz = s.dims
    assert (tmp_path / "simulation_check").read_text() == f"nx:{nx} ny:{ny} nz:{nz}\nTotal size:{nx * ny * nz}\n"

    def f(v):
        return "%f" % v

    def mat(name, eps_r, mu_r, se, sm):
        e0, m0 = 8.854187817e-12, 4e-7 * math.pi
        er, mr = float(np.float32(eps_r)), float(np.float32(mu_r))
        return (f"Material: {name}\neps_r: {f(er)}\nmu_r: {f(mr)}\neps: {f(er * e0)}\nmu: {f(mr * m0)}\n"
                f"sigma_e: {f(se)}\nsigma_m: {f(sm)}\nis_dispersion: 0")
    # the C API names the objects it builds by shape and their materials
    # "material" (xf_sim_add_cube / xf_sim_add_sphere)
    want = ("Object:\n"
            "Object: cube\nCube: " + " ".join(f(v) for v in (-0.05, -0.04, -0.03, 0.1, 0.08, 0.06)) + "\n"
            + mat("material", 1, 1, 0, 0) + "\n"
            "Object: sphere\nSphere: " + " ".join(f(v) for v in (0.01, 0.0, -0.005, 0.02)) + "\n"
            + mat("material", 2.25, 1.0, 0.01, 0.0) + "\n")
    assert (tmp_path / "simulation_object_check").read_text() == want
    s.close()
