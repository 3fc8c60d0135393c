"""Scenes shared by the adjoint tests (CPU and GPU)."""
import math

import numpy as np

from xfdtd_amd import scene as S


def box_scene(n=(14, 12, 13), dl=3e-3, pml=3, kappa=3.0):
    sc = S.Scene(cell_size=dl, cfl=0.95)
    size = tuple(v * dl for v in n)
    sc.add_object(S.Object("box", S.Cube(tuple(-s / 2 for s in size), size), S.Material.air()))
    sc.add_object(S.Object("c", S.Cube((-0.006, -0.005, -0.004), (0.012, 0.009, 0.01)),
                           S.Material("m", 3.0, 2.0, 0.2, 0.0)))
    sc.boundaries = [S.PML(o, pml, 3, 0.8, 0.0, 0.05, kappa) for o in range(6)]
    return sc


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


def drude_box_scene(n=(14, 12, 13), dl=3e-3, pml=3, kappa=3.0):
    """box_scene with a Drude cube added before the lossy cube, which then
    covers part of it (those cells: the cube's standard H coefficients, the
    Drude E update)."""
    sc = S.Scene(cell_size=dl, cfl=0.95)
    size = tuple(v * dl for v in n)
    sc.add_object(S.Object("box", S.Cube(tuple(-s / 2 for s in size), size), S.Material.air()))
    sc.add_object(S.Object("plasma", S.Cube((0.0, -0.006, -0.0045), (0.009, 0.009, 0.006)),
                           S.DrudeMedium("p", 1.5, 2 * math.pi * 30e9, 2 * math.pi * 2e9, 0.05, 0.0)))
    sc.add_object(S.Object("c", S.Cube((-0.006, -0.005, -0.004), (0.012, 0.009, 0.01)),
                           S.Material("m", 3.0, 2.0, 0.2, 0.0)))
    sc.boundaries = [S.PML(o, pml, 3, 0.8, 0.0, 0.05, kappa) for o in range(6)]
    return sc


def drude_design_scene(n=26, dl=4e-3, pml=5, steps=160, seed=0xC5):
    """design_scene with a Drude block between the design box and the
    objective plane."""
    sc, steps = design_scene(n, dl, pml, steps, seed)
    half = n * dl / 2
    sc.add_object(S.Object("plasma", S.Cube((-half + 15 * dl, -half + 10 * dl, -half + 11 * dl), (2 * dl, 5 * dl, 4 * dl)),
                           S.DrudeMedium("p", 1.0, 2 * math.pi * 20e9, 2 * math.pi * 1e9)))
    return sc, steps
