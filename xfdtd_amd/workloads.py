"""Seeded synthetic scatter workloads.

Random draws use a repository-owned splitmix64 generator with explicit
transforms, so a seed gives the same scenes on every platform and numpy
version.
"""
from __future__ import annotations

import math

import numpy as np

from . import scene as S

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): state += golden gamma, then two
    xor-shift-multiply rounds."""

    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        """53-bit uniform double in [lo, hi)."""
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def uniform_array(self, n, lo=0.0, hi=1.0):
        """The next n values of uniform(lo, hi), computed together (the same
        bits as n calls)."""
        g = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + g * np.arange(1, n + 1, dtype=np.uint64)
            self.state = int(z[-1]) if n else self.state
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi] (inclusive), by rejection."""
        span = hi - lo + 1
        limit = (1 << 64) - ((1 << 64) % span)
        while True:
            x = self.next_u64()
            if x < limit:
                return lo + x % span


def random_scatter_scene(rng, n, dl=5e-3, pml=8, tfsf_gap=12, cfl=0.9, nffft=True):
    """One batched-scatter instance: an n^3 total grid (CPML included) with
    1-3 dielectric spheres or boxes inside the total-field box and a plane
    wave at a random angle on the upper hemisphere.

    - objects: sphere or cube with probability 1/2 each; eps_r uniform in
      [1.5, 6] (rounded to float by the material model); radius / half-edge
      uniform in [0.05 n, 0.25 n] cells, capped so the object fits inside the
      TF box; centre uniform over the positions where it fits;
    - incidence: theta = arccos(u), u uniform in [0, 1]; phi uniform in
      [0, 2 pi); psi = 0;
    - NF-FF surface 2 cells inside the TF box at f = c0 / (20 dl).
    """
    n_obj = rng.integer(1, 3)
    half = (n - 2 * pml) * dl / 2
    tf_lo = -half + (tfsf_gap - pml) * dl      # TF box in coordinates
    tf_hi = half - (tfsf_gap - pml) * dl
    objs = []
    for q in range(n_obj):
        kind = "sphere" if rng.uniform() < 0.5 else "cube"
        eps = rng.uniform(1.5, 6.0)
        r_cells = rng.uniform(0.05 * n, 0.25 * n)
        r = min(r_cells * dl, 0.45 * (tf_hi - tf_lo))
        lo, hi = tf_lo + r, tf_hi - r
        c = tuple(rng.uniform(lo, hi) for _ in range(3))
        mat = S.Material(f"m{q}", eps, 1.0, 0.0, 0.0)
        if kind == "sphere":
            objs.append(S.Object(f"o{q}", S.Sphere(c, r), mat))
        else:
            objs.append(S.Object(f"o{q}", S.Cube(tuple(x - r for x in c), (2 * r,) * 3), mat))
    theta = math.acos(rng.uniform())
    phi = rng.uniform(0.0, 2 * math.pi)
    return S.cube_domain_scene(
        n, dl, pml, cfl=cfl, objects=objs, tfsf_gap=tfsf_gap,
        nffft_gap=(tfsf_gap - 2) if nffft else None,
        freqs=[S.C_0 / (20 * dl)], theta=[i * math.pi / 18 for i in range(19)], phi=[0.0],
        incidence=(theta, phi, 0.0))


def batched_scatter_set(count=64, seed=0xC3, n_min=48, n_max=160, **kw):
    """`count` instances with grid edges uniform in {n_min..n_max}."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = rng.integer(n_min, n_max)
        out.append(random_scatter_scene(rng, n, **kw))
    return out


def adjoint_design_scene(n=128, dl=4e-3, pml=10, src=(20, 64, 64), box=(44, 84), plane_x=108,
                         plane=(10, 118), seed=0xC4):
    """The design-gradient workload: an n^3 total grid (CPML included, order
    3, sigma 0.8 sigma_opt, kappa 1, alpha 0.05), cfl 0.99, a soft z current
    source (R -> infinity) at Ez(src) with a Gaussian of tau = 20 dt,
    t0 = 4.5 tau, a design cube [box)^3 of eps_r i.i.d. uniform in [1, 4]
    (splitmix64 `seed`, rounded to float) and the objective
    J = sum |DFT Ez(plane_x, j, k)|^2 over j, k in [plane) at f0 = c0/(20 dl)."""
    sc = S.cube_domain_scene(n, dl, pml, cfl=0.99)
    dt = S.courant_dt(dl, 0.99)
    tau = 20 * dt
    half = n * dl / 2
    org = tuple(-half + c * dl for c in src)
    sc.sources.append(S.CurrentSource(S.Cube(org, (dl, dl, dl)), S.ZP, math.inf,
                                      S.GaussianWaveform(tau, 4.5 * tau, 1.0)))
    rng = SplitMix64(seed)
    m = box[1] - box[0]
    eps = rng.uniform_array(m ** 3, 1.0, 4.0)
    sc.design = S.Design((box[0], box[1]) * 3, eps)
    sc.dft_plane = S.DftPlane(S.EZ, 0, plane_x, plane[0], plane[1], plane[0], plane[1], S.C_0 / (20 * dl))
    return sc



def batched_adjoint_set(count=32, seed=0xC5, n=96, dl=4e-3, pml=8, box=(32, 64), drude=(20, 30),
                        tfsf_gap=12):
    """The batched design-gradient workload: `count` n^3 total grids (CPML
    included, order 3, sigma 0.8 sigma_opt, kappa 1, alpha 0.05), cfl 0.9,
    each with

    - a Drude block of cells [drude)^3: eps_inf 1, omega_p 2 pi 20 GHz,
      nu 2 pi 1 GHz;
    - a design cube [box)^3 of eps_r i.i.d. uniform in [1, 4] (rounded to
      float);
    - a TFSF Gaussian plane wave (distance tfsf_gap) at theta = arccos(u),
      u uniform in [0, 1], phi uniform in [0, 2 pi), psi = 0;
    - the objective J = sum |DFT Ez(n - pml - 10, j, k)|^2 over j, k in
      [pml, n - pml) at f0 = c0/(20 dl).

    Per instance the splitmix64 stream gives theta, phi, then the design
    permittivities in C order.
    """
    rng = SplitMix64(seed)
    half = (n - 2 * pml) * dl / 2 + pml * dl
    out = []
    for _ in range(count):
        theta = math.acos(rng.uniform())
        phi = rng.uniform(0.0, 2 * math.pi)
        lo = -half + drude[0] * dl
        plasma = S.Object("plasma", S.Cube((lo, lo, lo), ((drude[1] - drude[0]) * dl,) * 3),
                          S.DrudeMedium("drude", 1.0, 2 * math.pi * 20e9, 2 * math.pi * 1e9))
        sc = S.cube_domain_scene(n, dl, pml, cfl=0.9, objects=[plasma], tfsf_gap=tfsf_gap,
                                 incidence=(theta, phi, 0.0))
        m = box[1] - box[0]
        sc.design = S.Design((box[0], box[1]) * 3, rng.uniform_array(m ** 3, 1.0, 4.0))
        sc.dft_plane = S.DftPlane(S.EZ, 0, n - pml - 10, pml, n - pml, pml, n - pml, S.C_0 / (20 * dl))
        out.append(sc)
    return out


# the 8 probe points of the 400^3 dielectric-sphere run, each sampled on Ez and Ex
SPHERE_400_PROBES = [(200, 200, 200), (200, 200, 310), (200, 200, 90), (120, 200, 200), (200, 330, 200),
                     (30, 200, 200), (6, 200, 200), (200, 200, 395)]


def pec_sphere_bistatic_scene(n=64, dl=5e-3, pml=8, radius_cells=16, cfl=0.9):
    """The bistatic PEC-sphere scatter run in the shape of the reference's
    examples/3d_pec_sphere_bistatic_rcs.cpp:34-101: an n^3 vacuum grid
    (CPML included, order 3, sigma 0.8 sigma_opt, kappa_max 1, alpha 0.05),
    a PEC sphere (sigma 1e10) of radius_cells at the centre, TFSF 4 cells
    inside the CPML with an x-polarised Gaussian plane wave along +z
    (tau = 20 dt, t0 = 4.5 tau), and the NF-FF box 2 cells outside the TFSF
    box at f = c0/(40 dl), c0/(30 dl), c0/(20 dl) over 181 theta in [0, pi]
    and phi in {0, pi/2}.  theta = pi is the monostatic direction."""
    objs = [S.Object("pec", S.Sphere((0.0, 0.0, 0.0), radius_cells * dl), S.Material.pec())]
    return S.cube_domain_scene(
        n, dl, pml, cfl=cfl, objects=objs, tfsf_gap=pml + 4, nffft_gap=pml + 2,
        freqs=[S.C_0 / (40 * dl), S.C_0 / (30 * dl), S.C_0 / (20 * dl)],
        theta=np.linspace(0.0, math.pi, 181), phi=[0.0, math.pi / 2])


def dielectric_sphere_scene(n=400, dl=2.5e-3, pml=12, radius_cells=100, probes=True):
    """The single-domain dielectric-sphere run: an n^3 grid (CPML included,
    order 3, sigma 0.8 sigma_opt, kappa_max 1, alpha 0.05), cfl 0.9, an
    eps_r = 4, sigma = 0 sphere of radius_cells at the centre, TFSF 4 cells
    inside the CPML (the same Gaussian plane wave as the PEC run), the NF-FF
    box 2 cells outside it at c0/(40 dl), c0/(30 dl), c0/(20 dl) over 181
    theta and phi in {0, pi/2}; at n = 400 Ez and Ex probes at the 8 points
    of SPHERE_400_PROBES.  For grids smaller than 400 the radius is scaled
    with n (at most n / 4) so the sphere stays clear of the CPML and the TFSF
    box."""
    if n != 400:
        radius_cells = max(2, min(radius_cells, n // 4))
    objs = [S.Object("sphere", S.Sphere((0.0, 0.0, 0.0), radius_cells * dl),
                     S.Material("sphere", 4.0, 1.0, 0.0, 0.0))]
    sc = S.cube_domain_scene(
        n, dl, pml, cfl=0.9, objects=objs, tfsf_gap=pml + 4, nffft_gap=pml + 2,
        freqs=[S.C_0 / (40 * dl), S.C_0 / (30 * dl), S.C_0 / (20 * dl)],
        theta=np.linspace(0.0, math.pi, 181), phi=[0.0, math.pi / 2])
    if probes and n == 400:
        sc.probes = [S.Probe(c, *p) for p in SPHERE_400_PROBES for c in (S.EZ, S.EX)]
    return sc
