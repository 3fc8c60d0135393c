"""Writes the golden fixtures of this directory: small scenes run by the
FP64 oracle (oracle/xfdtd_oracle.c), stored as compressed .npz files with
the scene parameters, the final fields and the scene's recorded series,
plus MANIFEST.txt with each file's sha256.

The reference holds no golden field vectors (its one known-answer test is
the grid test restated in test_oracle_cpu.py), so these fixtures freeze the
oracle's outputs: tests/test_golden_cpu.py checks that the oracle still
reproduces them bit for bit, and tests/test_golden_gpu.py checks the GPU
engine against them without building or running the oracle.

usage: python tests/golden/make_golden.py
"""
import hashlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))


def scenes():
    """name -> (scene builder, steps); the builders are deterministic."""
    from xfdtd_amd import scene as S
    from test_host_cpu import lumped_scene

    def sphere_tfsf():
        dl = 5e-3
        sph = S.Object("d", S.Sphere((1e-3, -2e-3, 0.5e-3), 3 * dl), S.Material("d", 3.0, 1.0, 0.01, 0.0))
        return S.cube_domain_scene(16, dl, 4, cfl=0.9, objects=[sph], tfsf_gap=5, nffft_gap=5,
                                   freqs=[2e9, 7e9], theta=[0.3, 2.0], phi=[0.0, 1.1],
                                   incidence=(0.7, 0.4, 0.3), tau_steps=6.0)

    def cylinder_2d():
        dl = 0.75e-3
        return S.cylinder_2d_scene(40, 36, 6, dl, 6.3 * dl, tfsf_gap=10, nffft_gap=8, phi_inc=0.3)

    def dispersive():
        # a Lorentz sphere, a lossy Debye cube, a Drude cube and a later plain
        # cube over part of the sphere, all clear of the 4-cell CPML
        dl, wp = 4e-3, 2 * math.pi * 20e9
        objs = [S.Object("lorentz", S.Sphere((1e-3, -2e-3, 0.0), 3 * dl),
                         S.LorentzMedium("l", 3.0, 1.5, wp, 2 * math.pi * 2e9)),
                S.Object("debye", S.Cube((-0.020, 0.004, -0.008), (0.008, 0.010, 0.008)),
                         S.DebyMedium("d", 5.0, 2.0, 1 / (2 * math.pi * 8e9), 0.03, 0.0)),
                S.Object("drude", S.Cube((0.008, 0.006, -0.016), (0.008, 0.008, 0.008)),
                         S.DrudeMedium("p", 1.0, 0.7 * wp, 2 * math.pi * 1e9)),
                S.Object("cap", S.Cube((0.002, -0.020, -0.020), (0.014, 0.040, 0.040)),
                         S.Material("c", 3.0, 1.2, 0.01, 0.0))]
        return S.cube_domain_scene(20, dl, 4, cfl=0.99, objects=objs, tfsf_gap=6, nffft_gap=5,
                                   freqs=[S.C_0 / (20 * dl)], theta=[0.0, 1.0], phi=[0.0])

    return {"sphere_tfsf": (sphere_tfsf, 40), "cylinder_2d": (cylinder_2d, 80),
            "lumped": (lumped_scene, 60), "dispersive": (dispersive, 40)}


def run_oracle(name):
    """The fixture's arrays from an oracle run."""
    from oracle.oracle import OracleSim
    build, steps = scenes()[name]
    sc = build()
    o = OracleSim(sc, steps)
    o.step(steps)
    out = {"steps": np.array(steps), "dt": np.array(o.dt), "dims": np.array(o.dims)}
    for c, n in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
        out[n] = o.get_field(c)
    if sc.nffft is not None and hasattr(sc.nffft, "frequencies"):
        out["nffft"] = o.get_nffft()
    if sc.nffft is not None and not hasattr(sc.nffft, "frequencies"):
        out["bb_samples"] = o.get_bb_samples()
    if any(type(m).__name__ in ("VoltageMonitor", "CurrentMonitor") for m in sc.monitors):
        out["line_monitors"] = o.get_line_monitors()
    if any(hasattr(ob.material, "eps_inf") for ob in sc.objects):
        for c, n in enumerate(("jx", "jy", "jz")):
            out[n] = o.get_drude_j(c)
    return out


def main():
    lines = []
    for name in scenes():
        data = run_oracle(name)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        h = hashlib.sha256(open(path, "rb").read()).hexdigest()
        lines.append(f"{h}  {name}.npz")
        print(name, os.path.getsize(path), "bytes")
    open(os.path.join(HERE, "MANIFEST.txt"), "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
