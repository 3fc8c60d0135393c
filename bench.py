"""Throughput of the field engine on one MI355X.

Workload: the 400^3 single-domain dielectric-sphere scatter run (dx = 2.5 mm,
12-cell CPML with order 3 / sigma 0.8 sigma_opt / kappa_max 1 / alpha 0.05,
eps_r = 4 sphere of radius 100 cells at the grid centre, x-polarised Gaussian
plane wave along +z through a TFSF box 16 cells from the edge, NF-FF surface
14 cells from the edge at 3 frequencies), FP32 fields.  A step is one full
time step of the reference's run loop: E curl + CPML, TFSF E faces, H curl +
CPML, 1-D incident grid, TFSF H faces, NF-FF running DFT.

Prints one JSON line.  value = N^3 * steps / wall time of the timed steps
(Gcell-updates/s), inputs resident in HBM.  roofline: algorithmic bytes of
the dominant curl kernel (18*s/2 + 1 bytes per cell per half step: read 3 E +
3 H, write 3 of one kind, one material byte) over its average HIP-event
duration inside the timed region; traffic: the HBM bytes per launch of that
kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by the same
counters' reading for a plain 1 GiB copy (profiles/hbm_traffic.json).
cpu_baseline: the FP64 CPU oracle, one thread, on six steps of the same 400^3
run (set-up excluded); cpu_baseline_threads: the same with every host core.

batched: 64 seeded scatter instances of uneven size (edges uniform in
{48..160}, 8-cell CPML, 1-3 random dielectric objects, TFSF at a random
angle, NF-FF at one frequency; xfdtd_amd/workloads.py), FP32, tile shapes from
the autotuner per size class.  The plain loop steps each instance alone and
waits for it before the next, as one Simulation at a time runs; the batched
run steps all of them together (xf_batch_step).  Both report
sum(N^3) * steps / wall time.

adjoint: the design-gradient workload (128^3 with 10-cell CPML, soft Ez
source, 40^3 design cube of random eps_r, J = sum |DFT Ez|^2 on the plane
x = 108; xfdtd_amd/workloads.py), 1500 steps, checkpoint every 100.
cost_ratio = wall time of the gradient (forward with checkpoints + segment
recompute + adjoint) / wall time of the plain forward run, per precision;
the FP64 gradient is checked against central differences (delta eps_r =
1e-3) at 16 seeded design cells, and the FP32 gradient against the FP64 one.

batched_adjoint: 32 gradients of 96^3 instances with a Drude block, a design
cube and TFSF at random incidence (xfdtd_amd/workloads.py), 1200 steps, as
one batch against a loop of single gradients.

broadband: the reference's monostatic example (96^3 PEC sphere, broadband
NF-FF, FP64, 1200 steps): stepping with and without the broadband surface,
and the far field of all 2T-1 bins (hipFFT + fold).

dispersive: the 400^3 run with its sphere made of a plain dielectric, then
Drude, Debye and Lorentz media (the E kernel's general path), 100 steps.

tmz: the reference's 2-D example (370 x 320 TMz cells, PEC cylinder, TFSF2D,
broadband NF-FF, 1400 steps) in FP64 and FP32, its far field, and the
oracle's step time on the same scene.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Yee Gcells/s & % HBM peak; batched vs plain; adjoint grad time at stated tol"


def sphere_radius(n):
    """The sphere radius (cells) of dielectric_sphere_scene at grid edge n."""
    return 100 if n == 400 else max(2, min(100, n // 4))


def scene_400(n=400, steps_total=2000):
    from xfdtd_amd.workloads import dielectric_sphere_scene
    return dielectric_sphere_scene(n), steps_total


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_cpu_baseline(sc, total, steps, threads=1):
    from oracle.oracle import OracleSim, set_threads
    threads = set_threads(threads)
    o = OracleSim(sc, total)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    nx, ny, nz = o.dims
    o.close()
    cells = nx * ny * nz
    return {
        "value": cells * steps / dt / 1e9,
        "unit": "Gcells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, {nx}x{ny}x{nz} grid of the same run, first {steps} steps "
                  f"timed after set-up ({dt:.1f} s), {threads} thread(s) on '{cpu_model()}' "
                  f"(nproc {os.cpu_count()})",
    }


def pattern_ceiling(n):
    """The E curl kernel's access pattern without its arithmetic, CPML, TFSF
    or tables (xfdtd_amd/pattern_model, built from csrc/pattern_model.hip):
    the same fields, layout, tiles and loads and stores per cell, timed on
    this GPU in FP32 and FP64.  Its algorithmic GB/s is the practical ceiling
    of the split curl kernels."""
    import subprocess
    exe = os.path.join(ROOT, "xfdtd_amd", "pattern_model")
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} is not built (make -C xfdtd_amd/csrc)")
    out = subprocess.run([exe, str(n)], check=True, capture_output=True, text=True, timeout=300).stdout
    res = {}
    for ln in out.splitlines():
        if ln.startswith("{"):
            d = json.loads(ln)
            res[d.pop("precision")] = d
    res["pattern"] = ("64 k x 4 j threads marching 8 x planes; per cell read Ex Ey Ez Hx Hy Hz, the material "
                      "byte, Hy Hx at k-1, Hz Hx at j-1, write Ex Ey Ez in place")
    return res


def measured_traffic(kernel):
    """HBM bytes per launch of a curl kernel from the committed counter runs."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][kernel]["bytes"], d
    except (OSError, KeyError, ValueError):
        return None, None


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size: its FETCH_SIZE/WRITE_SIZE
    counters calibrate the counter readings of the field kernels."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": nbytes, "copy_ms": dt * 1e3,
                      "copy_GBps_read_plus_write": 2 * nbytes / dt / 1e9}), flush=True)


def run_batched(count, steps, warmup=10, log=None):
    import numpy as np

    from xfdtd_amd import autotune
    from xfdtd_amd.sim import Batch, Simulation
    from xfdtd_amd.workloads import batched_scatter_set
    table = autotune.tune(precision=32, log=log)
    scenes = batched_scatter_set(count)
    total = warmup + steps
    tune_steps = 5 + 40

    def build(tiling=None, horizon=total):
        sims = []
        for sc in scenes:
            sim = Simulation(sc, horizon, precision=32)
            sim.set_tiling(*(tiling or autotune.tiling_for(table, sim.dims[0])))
            sims.append(sim)
        return sims

    # the batch's own tiling, timed on the batched launch (throw-away members)
    probe = build(horizon=tune_steps * len(autotune.BATCH_CANDIDATES))
    batch_tiling = autotune.tune_batch(probe, log=log)
    for sim in probe:
        sim.close()

    cells = 0
    plain = build()
    for sim in plain:
        nx, ny, nz = sim.dims
        cells += nx * ny * nz
        sim.step(warmup)
        sim.sync()
    t0 = time.perf_counter()
    for sim in plain:
        sim.step(steps)
        sim.sync()
    t_plain = time.perf_counter() - t0
    batch_sims = build(batch_tiling)
    b = Batch(batch_sims)
    b.step(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    # the batched run must reproduce the plain one exactly (same kernels)
    same = all(np.array_equal(plain[q].get_field(c), batch_sims[q].get_field(c))
               for q in (0, count - 1) for c in range(6))
    b.close()
    for sim in plain + batch_sims:
        sim.close()
    return {
        "workload": f"{count} seeded scatter instances (splitmix64 seed 0xC3), grid edges uniform in "
                    "{48..160} incl. 8-cell CPML, 1-3 dielectric spheres/boxes, TFSF at random incidence, "
                    "NF-FF 1 freq, FP32",
        "instances": count,
        "cells_per_step": cells,
        "steps": steps,
        "plain_loop_gcells": round(cells * steps / t_plain / 1e9, 3),
        "batched_gcells": round(cells * steps / t_batch / 1e9, 3),
        "speedup": round(t_plain / t_batch, 3),
        "bit_identical": bool(same),
        "tilings": {f"{lo}-{hi}": list(t) for (lo, hi), t in table.items()},
        "batch_tiling": list(batch_tiling),
    }


def run_adjoint(fd_cells=16, steps=1500, every=100):
    import numpy as np

    from xfdtd_amd import scene as S
    from xfdtd_amd.sim import Simulation
    from xfdtd_amd.workloads import SplitMix64, adjoint_design_scene
    sc = adjoint_design_scene()
    out = {"workload": "128^3 incl. 10-cell CPML, dx 4 mm, cfl 0.99, soft Jz source at (20,64,64), design "
                       "cube [44,84)^3 eps_r ~ U[1,4] (splitmix64 seed 0xC4), J = sum |DFT Ez(108,j,k)|^2, "
                       "j,k in [10,118), f0 = c0/(20 dx)",
           "steps": steps, "checkpoint_every": every}
    grads = {}
    for prec in (32, 64):
        fwd = Simulation(sc, steps, precision=prec)
        fwd.sync()
        t0 = time.perf_counter()
        fwd.step(steps)
        fwd.sync()
        t_fwd = time.perf_counter() - t0
        fwd.close()
        gs = Simulation(sc, steps, precision=prec)
        gs.sync()
        t0 = time.perf_counter()
        J, g = gs.grad(every)
        t_grad = time.perf_counter() - t0
        gs.close()
        # one segment: the design history of the whole run is kept during the
        # forward run, nothing is recomputed
        gs = Simulation(sc, steps, precision=prec)
        gs.sync()
        t0 = time.perf_counter()
        _, g1 = gs.grad(steps)
        t_one = time.perf_counter() - t0
        gs.close()
        grads[prec] = g
        out[f"fp{prec}"] = {"forward_s": round(t_fwd, 4), "grad_s": round(t_grad, 4),
                            "cost_ratio": round(t_grad / t_fwd, 3),
                            "grad_no_recompute_s": round(t_one, 4),
                            "cost_ratio_no_recompute": round(t_one / t_fwd, 3),
                            "no_recompute_bit_identical": bool(np.array_equal(g, g1)), "J": J}
    g64, g32 = grads[64].ravel(), grads[32].ravel()
    out["fp32_vs_fp64_grad_rel"] = float(np.abs(g32 - g64).max() / np.abs(g64).max())
    if fd_cells > 0:
        rng = SplitMix64(0xC4F)
        cells = []
        while len(cells) < fd_cells:
            c = rng.integer(0, g64.size - 1)
            if c not in cells:
                cells.append(c)
        eps = np.asarray(sc.design.eps_r, dtype=np.float32)
        errs = []
        for c in cells:
            vals = []
            hi, lo = np.float32(float(eps[c]) + 1e-3), np.float32(float(eps[c]) - 1e-3)
            for v in (hi, lo):
                e = eps.copy()
                e[c] = v
                sc.design = S.Design(sc.design.box, e)
                sim = Simulation(sc, steps, precision=64)
                sim.step(steps)
                vals.append(float(np.sum(np.abs(sim.get_dft()) ** 2)))
                sim.close()
            fd = (vals[0] - vals[1]) / (float(hi) - float(lo))
            errs.append(abs(g64[c] - fd) / max(abs(fd), 1e-3 * np.abs(g64).max()))
        sc.design = S.Design(sc.design.box, eps)
        out["fd_cells"] = fd_cells
        out["fp64_grad_vs_central_differences_max_rel"] = float(max(errs))
    return out


def run_batched_adjoint(count=32, steps=1200, every=0):
    """Batched design gradients: `count` 96^3 instances with a Drude block,
    a 32^3 design cube and a TFSF source each, stepped and differentiated as
    one batch (FP32), against a loop of single gradients and the batched
    forward run alone."""
    import numpy as np

    from xfdtd_amd.sim import Batch, Simulation
    from xfdtd_amd.workloads import batched_adjoint_set
    scenes = batched_adjoint_set(count)

    def build(prec=32):
        return [Simulation(sc, steps, precision=prec) for sc in scenes]

    sims = build()
    cells = sum(s.dims[0] * s.dims[1] * s.dims[2] for s in sims)
    b = Batch(sims)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_fwd = time.perf_counter() - t0
    b.close()
    for s in sims:
        s.close()
    sims = build()
    b = Batch(sims)
    b.sync()
    t0 = time.perf_counter()
    batched = b.grad(every)
    t_grad = time.perf_counter() - t0
    b.close()
    for s in sims:
        s.close()
    t0 = time.perf_counter()
    loop = []
    for sc in scenes:
        s = Simulation(sc, steps, precision=32)
        s.sync()
        loop.append(s.grad(every))
        s.close()
    t_loop = time.perf_counter() - t0
    same = all(a[0] == c[0] and np.array_equal(a[1], c[1]) for a, c in zip(batched, loop))
    ref = Simulation(scenes[0], steps, precision=64)
    _, g64 = ref.grad(every)
    ref.close()
    g32 = batched[0][1]
    return {
        "workload": f"{count} instances of 96^3 incl. 8-cell CPML, dx 4 mm, cfl 0.9, Drude block [20,30)^3 "
                    "(eps_inf 1, omega_p 2pi 20 GHz, nu 2pi 1 GHz), design cube [32,64)^3 eps_r ~ U[1,4], "
                    "TFSF at random incidence (splitmix64 seed 0xC5), J = sum |DFT Ez(78,j,k)|^2, j,k in [8,88), "
                    "f0 = c0/(20 dx), FP32",
        "instances": count,
        "steps": steps,
        "checkpoint_every": every if every else "auto",
        "cells_per_step": cells,
        "batched_forward_s": round(t_fwd, 4),
        "batched_forward_gcells": round(cells * steps / t_fwd / 1e9, 3),
        "batched_grad_s": round(t_grad, 4),
        "cost_ratio": round(t_grad / t_fwd, 3),
        "loop_of_single_grads_s": round(t_loop, 4),
        "speedup_vs_loop": round(t_loop / t_grad, 3),
        "bit_identical_to_single": bool(same),
        "member0_fp32_vs_fp64_grad_rel": float(np.abs(g32 - g64).max() / np.abs(g64).max()),
    }


def monostatic_scene(broadband=True):
    """The reference's monostatic example (examples/3d_pec_sphere_monostatic_rcs.cpp):
    1.6 m air domain at dl = 2 cm with 8-cell CPML (96^3), PEC sphere of
    radius 0.5 m, TFSF at 15 along +z, broadband NF-FF box at 12 observed
    along -z, cfl 0.98."""
    import math

    from xfdtd_amd import scene as S
    dl = 2e-2
    sc = S.Scene(cell_size=dl, cfl=0.98)
    sc.add_object(S.Object("domain", S.Cube((-0.8,) * 3, (1.6,) * 3), S.Material("air", 1, 1, 0, 0)))
    sc.add_object(S.Object("scatter", S.Sphere((0, 0, 0), 0.5), S.Material("pec", 1, 1, 1e10, 0)))
    sc.boundaries = [S.PML(o, 8) for o in range(6)]
    tau = 20 * dl / 6e8
    sc.tfsf = S.TFSF3D(15, 15, 15, 1.0, 0.0, 0.0, 0.0, S.GaussianWaveform(tau, 4.5 * tau, 1.0))
    if broadband:
        sc.nffft = S.NffftBroadBand(12, 12, 12, math.pi, math.pi)
    return sc


def run_broadband(steps=1200):
    """The monostatic example in FP64: stepping with and without the
    broadband surface, and the far field of all 2T-1 bins."""
    from xfdtd_amd.sim import Simulation
    out = {"workload": "reference monostatic example: 96^3 incl. 8-cell CPML, dl 2 cm, PEC sphere r 0.5 m, "
                       "TFSF Gaussian along +z, broadband NF-FF box at 12 observed along -z, FP64",
           "steps": steps, "bins": 2 * steps - 1}
    for key, bb in (("step_s_without_surface", False), ("step_s", True)):
        sim = Simulation(monostatic_scene(bb), steps, precision=64)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        out[key] = round(time.perf_counter() - t0, 4)
        if bb:
            sim.bb_farfield()  # hipFFT plans and kernels warm
            t0 = time.perf_counter()
            sim.bb_farfield()
            out["far_field_s"] = round(time.perf_counter() - t0, 4)
            out["series"] = int(sim.get_bb_samples().shape[2]) * 8
        sim.close()
    return out


def run_dispersive(n=400, steps=100, warmup=5):
    """The headline 400^3 run (FP32) with its sphere (radius 100 cells at
    400^3, scaled with smaller grids) made of each medium in turn: a plain eps_r 4 dielectric, a Drude plasma (eps_inf 1,
    omega_p 2 pi 20 GHz, nu 2 pi 1 GHz), a Debye medium (eps_s 4, eps_inf 2,
    tau 1 / (2 pi 5 GHz)) and a Lorentz medium (eps_s 4, eps_inf 2, omega_p
    2 pi 15 GHz, nu 2 pi 1 GHz).  The dispersive cells take
    the E kernel's general path; the figures are whole-step Gcells/s and the
    E and H kernels' HIP-event times."""
    from xfdtd_amd import scene as S
    from xfdtd_amd.sim import Simulation
    media = {"plain": S.Material("sphere", 4.0, 1.0, 0.0, 0.0),
             "drude": S.DrudeMedium("p", 1.0, 2 * math.pi * 20e9, 2 * math.pi * 1e9),
             "debye": S.DebyMedium("d", 4.0, 2.0, 1 / (2 * math.pi * 5e9)),
             "lorentz": S.LorentzMedium("l", 4.0, 2.0, 2 * math.pi * 15e9, 2 * math.pi * 1e9)}
    out = {"workload": f"{n}^3 headline scene (FP32), sphere of radius {sphere_radius(n)} cells made of each medium",
           "steps": steps}
    for name, mat in media.items():
        sc, _ = scene_400(n=n)
        sc.objects[-1] = S.Object("sphere", sc.objects[-1].shape, mat)
        sim = Simulation(sc, warmup + steps, precision=32)
        sim.step(warmup)
        sim.sync()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        dt = time.perf_counter() - t0
        e_ms, h_ms, _ = sim.kernel_times()
        nx, ny, nz = sim.dims
        out[name] = {"gcells": round(nx * ny * nz * steps / dt / 1e9, 2), "e_ms": round(e_ms, 4),
                     "h_ms": round(h_ms, 4)}
        sim.close()
    return out


def run_tmz(steps=1400, cpu_steps=200, cpu_threads=16):
    """The reference's 2-D example (examples/2d_cylinder_scatter.cpp): a
    370 x 320 TMz grid (350 x 300 plus a 10-cell CPML on x and y) with a PEC
    cylinder, TFSF2D along pi/4 and a broadband NF-FF box, FP64 and FP32
    steps, the far field of all 2T-1 bins, and the oracle's step time on the
    same scene (CPU, cpu_threads threads, cpu_steps steps)."""
    from xfdtd_amd import scene as S
    from xfdtd_amd.sim import Simulation
    dx = 3e8 / 20e9 / 20
    sc = S.cylinder_2d_scene(350, 300, 10, dx, 0.03, tfsf_gap=50, nffft_gap=40)
    out = {"workload": "reference 2-D example: 370 x 320 TMz cells incl. 10-cell CPML on x and y, PEC cylinder "
                       "r 3 cm, TFSF2D along pi/4, broadband NF-FF box at 40",
           "steps": steps, "cells": 370 * 320}
    for prec in (64, 32):
        sim = Simulation(sc, steps, precision=prec)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        dt = time.perf_counter() - t0
        out[f"fp{prec}_step_s"] = round(dt, 4)
        out[f"fp{prec}_us_per_step"] = round(1e6 * dt / steps, 2)
        out[f"fp{prec}_mcells_per_s"] = round(370 * 320 * steps / dt / 1e6, 1)
        if prec == 64:
            sim.bb_farfield()
            t0 = time.perf_counter()
            sim.bb_farfield()
            out["far_field_s"] = round(time.perf_counter() - t0, 4)
        sim.close()
    if cpu_steps > 0:
        # the leg's CPU baseline: the FP64 oracle on the same scene
        from oracle.oracle import OracleSim, set_threads
        threads = set_threads(cpu_threads)
        o = OracleSim(sc, cpu_steps)
        t0 = time.perf_counter()
        o.step(cpu_steps)
        cdt = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": round(1e6 * cdt / cpu_steps, 1), "unit": "us per step", "cores": threads,
                               "kind": "port", "sample": f"FP64 oracle, first {cpu_steps} steps of the same scene"}
    return out


def timed_run(sc, total, precision, steps, warmup, tiling=None, sweep=None):
    """Warm-up steps, then `steps` timed steps of one Simulation with the
    curl kernels' HIP events recorded on the engine's stream."""
    from xfdtd_amd.sim import Simulation
    sim = Simulation(sc, total, precision=precision)
    if tiling:
        sim.set_tiling(*tiling)
    if sweep:
        sim.set_sweep(*sweep)
    nx, ny, nz = sim.dims
    fused = sim.sweep_active
    sim.step(warmup)
    sim.sync()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ms_e, ms_h, nsamp = sim.kernel_times()
    sim.set_timing(False)
    sim.close()
    return {"grid": [nx, ny, nz], "wall_s": wall, "gcells": nx * ny * nz * steps / wall / 1e9,
            "ms_per_step": wall / steps * 1e3, "ms_e": ms_e, "ms_h": ms_h, "event_samples": nsamp,
            "fused": fused}


def dominant(run, s):
    """The dominant field kernel of a timed run, its average HIP-event time and
    its algorithmic bytes per launch per cell: the fused step (k_sweep) does a
    whole cell-step, 18 s + 2 bytes (read 3 E + 3 H and write 3 E, read 3 H
    + 3 E and write 3 H, one material byte per half step); a split kernel does
    a half step, 9 s + 1 bytes."""
    if run["fused"]:
        return "k_sweep", run["ms_e"], 18 * s + 2
    if run["ms_h"] >= run["ms_e"]:
        return "k_update_h", run["ms_h"], 9 * s + 1
    return "k_update_e", run["ms_e"], 9 * s + 1


def fp64_leg(sc, total, steps, warmup, tiling=None):
    """The same workload in FP64 (the reference's precision)."""
    r = timed_run(sc, total, 64, steps, warmup, tiling)
    nx, ny, nz = r["grid"]
    cells = nx * ny * nz
    kname, ms_dom, bpc = dominant(r, 8)
    achieved = bpc * cells / (ms_dom * 1e-3) / 1e9
    return {"value": round(r["gcells"], 3), "unit": "Gcells/s", "steps": steps, "warmup": warmup,
            "ms_per_step": round(r["ms_per_step"], 4),
            "avg_ms": {kname: round(ms_dom, 4)} if r["fused"] else
                      {"k_update_e": round(r["ms_e"], 4), "k_update_h": round(r["ms_h"], 4)},
            "roofline": {"kernel": kname, "bytes_per_launch": bpc * cells, "achieved": round(achieved, 1),
                         "frac": round(achieved / HBM_PEAK_GBS, 4)}}


def fused_leg(sc, total, steps, warmup, precision):
    """The same workload through the fused E+H kernel (time-step scheme 1,
    k_sweep), for comparison with the split kernels of the headline line."""
    r = timed_run(sc, total, precision, steps, warmup, sweep=(1, 2, 16))
    if not r["fused"]:
        return {"error": "the scene does not take the fused kernel"}
    nx, ny, nz = r["grid"]
    s = 4 if precision == 32 else 8
    kname, ms_dom, bpc = dominant(r, s)
    achieved = bpc * nx * ny * nz / (ms_dom * 1e-3) / 1e9
    return {"value": round(r["gcells"], 3), "unit": "Gcells/s", "steps": steps, "warmup": warmup,
            "ms_per_step": round(r["ms_per_step"], 4), "avg_ms": {kname: round(ms_dom, 4)},
            "geometry": {"rows": 2, "planes": 16},
            "roofline": {"kernel": kname, "bytes_per_launch": bpc * nx * ny * nz, "achieved": round(achieved, 1),
                         "frac": round(achieved / HBM_PEAK_GBS, 4)}}


def run_fp32_vs_fp64(n=400, steps=2000):
    """The whole run in FP64 and in FP32: the FP32 error at the probe points
    (max over time of |FP32 - FP64| over the series' peak, Ez and Ex at 8
    points), on the NF-FF far-field power per frequency (relative to that
    frequency's peak over all directions, and at the monostatic direction)
    and on the final fields (over each component's final peak)."""
    import numpy as np

    from xfdtd_amd.sim import Simulation
    from xfdtd_amd.workloads import SPHERE_400_PROBES, dielectric_sphere_scene
    sc = dielectric_sphere_scene(n)
    runs = {}
    for prec in (64, 32):
        s = Simulation(sc, steps, precision=prec)
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        runs[prec] = (s, time.perf_counter() - t0)
    s64, s32 = runs[64][0], runs[32][0]
    out = {"steps": steps, "fp64_run_s": round(runs[64][1], 3), "fp32_run_s": round(runs[32][1], 3)}
    if sc.probes:
        # per point: max over time of |FP32 - FP64| on Ez and Ex over the
        # point's field peak (the larger of its two series' peaks)
        p64, p32 = s64.get_probes(), s32.get_probes()
        perr = {}
        for q in range(0, len(p64), 2):
            err = max(np.abs(p32[q] - p64[q]).max(), np.abs(p32[q + 1] - p64[q + 1]).max())
            peak = max(np.abs(p64[q]).max(), np.abs(p64[q + 1]).max())
            perr[str(tuple(SPHERE_400_PROBES[q // 2]))] = float(err / peak)
        out["probes"] = perr
        out["probe_max_rel"] = max(perr.values())
    f64, f32 = s64.farfield(), s32.farfield()
    pw64 = f64["power_theta"] + f64["power_phi"]
    pw32 = f32["power_theta"] + f32["power_phi"]
    out["far_power_max_rel"] = [float(np.abs(pw32[q] - pw64[q]).max() / np.abs(pw64[q]).max())
                                for q in range(pw64.shape[0])]
    out["monostatic_power_rel"] = [float(abs(pw32[q, -1, 0] - pw64[q, -1, 0]) / abs(pw64[q, -1, 0]))
                                   for q in range(pw64.shape[0])]
    fld = []
    for c in range(6):
        a, b = s32.get_field(c), s64.get_field(c)
        fld.append(float(np.abs(a - b).max() / np.abs(b).max()))
    out["final_field_max_rel"] = fld
    s64.close()
    s32.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--size", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the second, multi-core CPU column (0: skip it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tiling", type=str, default="")
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=1500)
    ap.add_argument("--adjoint", type=int, default=1, help="run the design-gradient leg (0: skip it)")
    ap.add_argument("--adjoint-fd-cells", type=int, default=16)
    ap.add_argument("--batch-adjoint", type=int, default=32,
                    help="instances of the batched-gradient leg (0: skip it)")
    ap.add_argument("--batch-adjoint-steps", type=int, default=1200)
    ap.add_argument("--broadband", type=int, default=1200,
                    help="steps of the monostatic broadband NF-FF leg (0: skip it)")
    ap.add_argument("--dispersive", type=int, default=100,
                    help="steps of the dispersive-media leg at --size (0: skip it)")
    ap.add_argument("--tmz", type=int, default=1400,
                    help="steps of the 2-D (TMz) cylinder leg (0: skip it)")
    ap.add_argument("--fused-steps", type=int, default=100,
                    help="timed steps of the same workload through the fused E+H kernel (0: skip it)")
    ap.add_argument("--fp64-steps", type=int, default=100,
                    help="timed steps of the FP64 run of the same workload (0: skip it)")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="steps of the FP32-against-FP64 accuracy runs (0: skip them)")
    ap.add_argument("--pattern", type=int, default=1,
                    help="time the curl kernels' access pattern alone (0: skip it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain 1 GiB device copy (reference for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is not built.", file=sys.stderr)
        sys.exit(2)

    sc, total = scene_400(n=args.size)
    if args.warmup + args.steps > total:
        total = args.warmup + args.steps
    tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else None
    run = timed_run(sc, total, args.precision, args.steps, args.warmup, tiling)
    nx, ny, nz = run["grid"]
    cells = nx * ny * nz
    value = run["gcells"]
    wall = run["wall_s"]
    ms_e, ms_h, nsamp = run["ms_e"], run["ms_h"], run["event_samples"]
    s = 4 if args.precision == 32 else 8
    dom, ms_dom, bytes_cell = dominant(run, s)
    achieved = bytes_cell * cells / (ms_dom * 1e-3) / 1e9
    traffic, tinfo = (None, None)
    if args.precision == 32 and args.size == 400 and not args.tiling:
        traffic, tinfo = measured_traffic(dom)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain single-domain dielectric sphere {nx}x{ny}x{nz}, dx=2.5mm, "
                        f"12-cell CPML, eps_r=4 sphere r={sphere_radius(args.size)} cells, TFSF Gaussian plane "
                        "wave, NF-FF 3 freqs",
            "grid": [nx, ny, nz],
            "precision": args.precision,
            "total_steps_of_run": total,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic / 1e9, 4),
            "traffic_unit": "GB per launch",
            "traffic_source": None if tinfo is None else tinfo.get("source"),
            "bytes_per_launch": bytes_cell * cells,
            "avg_ms": {dom: round(ms_dom, 4)} if run["fused"] else
                      {"k_update_e": round(ms_e, 4), "k_update_h": round(ms_h, 4)},
            "event_samples": nsamp,
        },
        "cpu_baseline": None,
    }
    legs = []
    if args.pattern:
        legs.append(("pattern_ceiling", lambda: pattern_ceiling(args.size)))
    if args.fused_steps > 0:
        legs.append(("fused", lambda: fused_leg(sc, total, args.fused_steps, args.warmup, args.precision)))
    if args.fp64_steps > 0 and args.precision == 32:
        legs.append(("fp64", lambda: fp64_leg(sc, total, args.fp64_steps, args.warmup, tiling)))
    if args.accuracy_steps > 0:
        legs.append(("fp32_vs_fp64", lambda: run_fp32_vs_fp64(args.size, args.accuracy_steps)))
    if args.adjoint:
        legs.append(("adjoint", lambda: run_adjoint(args.adjoint_fd_cells)))
    if args.batch_adjoint > 0:
        legs.append(("batched_adjoint", lambda: run_batched_adjoint(args.batch_adjoint, args.batch_adjoint_steps)))
    if args.broadband > 0:
        legs.append(("broadband", lambda: run_broadband(args.broadband)))
    if args.dispersive > 0:
        legs.append(("dispersive", lambda: run_dispersive(args.size, args.dispersive)))
    if args.tmz > 0:
        legs.append(("tmz", lambda: run_tmz(args.tmz, 0 if args.no_cpu else 200, max(1, args.cpu_threads))))
    if args.batch > 0:
        legs.append(("batched", lambda: run_batched(args.batch, args.batch_steps,
                                                    log=lambda m: print(m, file=sys.stderr, flush=True))))
    if not args.no_cpu and args.cpu_steps > 0:
        legs.append(("cpu_baseline", lambda: run_cpu_baseline(sc, total, args.cpu_steps)))
        if args.cpu_threads > 1:
            legs.append(("cpu_baseline_threads",
                         lambda: run_cpu_baseline(sc, total, args.cpu_steps, args.cpu_threads)))
    for name, fn in legs:
        # a leg that fails records its error; the line is printed regardless
        print(f"bench: {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            line[name] = fn()
        except Exception as exc:  # noqa: BLE001
            line[name] = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"bench: {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    pc = line.get("pattern_ceiling") or {}
    key = "f32" if args.precision == 32 else "f64"
    if key in pc:
        ceil_gbs = pc[key]["algorithmic_gbs"]
        line["roofline"]["pattern_ceiling"] = ceil_gbs
        line["roofline"]["frac_of_pattern"] = round(achieved / ceil_gbs, 4)
    f64 = line.get("fp64")
    if "f64" in pc and isinstance(f64, dict) and "roofline" in f64:
        ceil_gbs = pc["f64"]["algorithmic_gbs"]
        f64["roofline"]["pattern_ceiling"] = ceil_gbs
        f64["roofline"]["frac_of_pattern"] = round(f64["roofline"]["achieved"] / ceil_gbs, 4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
