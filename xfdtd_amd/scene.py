"""Scene description mirroring the reference's C++ construction API.

Each class carries the constructor arguments of its reference counterpart:

  Cube(origin, size)                      shape/cube.cpp:54-55
  Sphere(center, radius)                  shape/sphere.cpp:14-15
  Material(name, eps_r, mu_r, sigma_e, sigma_m)
                                          material/material.cpp:7-16 (eps_r, mu_r are float)
  Object(name, shape, material)           object/object.cpp:18-26
  PML(orientation, thickness, order=4, sigma_ratio=1, alpha_min=0,
      alpha_max=0.05, kappa_max=10)       include/boundary/perfect_match_layer.h:33-35
  GaussianWaveform(amplitude, tau, t_0)   waveform/gaussian_waveform.cpp:17-21
  TFSF3D(distance_x, distance_y, distance_z, e_0, theta, phi, psi, waveform)
                                          tfsf/tfsf_3d.cpp:13-31
  NffftFd(distance_x, distance_y, distance_z, frequencies, theta, phi)
                                          include/nffft/nffft_fd.h (constructor)
  NffftBroadBand(distance_x, distance_y, distance_z, theta, phi)
                                          nffft/nffft_broadband.cpp:21-28
  CurrentSource(cube, orientation, resistance, waveform)
                                          lumped_element/current_source.cpp:5-16
  VoltageSource(cube, orientation, resistance, waveform)
                                          lumped_element/voltage_source.cpp:12-22
  Resistor / Capacitor / Inductor(cube, axis, value)
                                          lumped_element/resistor.cpp:7-8, capacitor.cpp:4-6,
                                          inductor.cpp:4-5
  ObjectPlane(name, shape, material)      object/object_plane.cpp:4-6
  VoltageMonitor / CurrentMonitor(cube, orientation, output_dir, output_file_name)
                                          monitor/voltage_monitor.cpp:11-18,
                                          monitor/current_monitor.cpp:11-18
  SinusoidalWaveform(amplitude, frequency, phase)
                                          waveform/sinusoidal_waveform.cpp:6-8
  CosineModulatedGaussianWaveform(amplitude, tau, t_0, modulation_frequency)
                                          waveform/cosine_modulated_gaussian_waveform.cpp:6-11
  Port(index, impedance, is_source, voltage, current), Network(ports, frequencies, output_path)
                                          network/port.cpp:6-13, network/network.cpp:8-11
  Simulation(cell_size, cfl=0.99f)        include/simulation/simulation.h:25-27

A `Scene` bundles them the way `Simulation::add*` does.  Both the HIP engine
(through the C-ABI) and the CPU oracle read this one description.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

# include/util/type_define.h:14
XN, XP, YN, YP, ZN, ZP = range(6)
ORIENTATIONS = {"XN": XN, "XP": XP, "YN": YN, "YP": YP, "ZN": ZN, "ZP": ZP}
EX, EY, EZ, HX, HY, HZ = range(6)
COMPONENTS = {"EX": EX, "EY": EY, "EZ": EZ, "HX": HX, "HY": HY, "HZ": HZ}

# util/constant.h
PI = math.pi
EPSILON_0 = 8.854187817e-12
MU_0 = 4 * PI * 1e-7
C_0 = math.sqrt(1 / (EPSILON_0 * MU_0))
ETA_0 = math.sqrt(MU_0 / EPSILON_0)


@dataclass
class Cube:
    origin: Tuple[float, float, float]
    size: Tuple[float, float, float]


@dataclass
class Sphere:
    center: Tuple[float, float, float]
    radius: float


@dataclass
class Cylinder:
    """Cylinder(axis, center, radius, height) (shape/cylinder.cpp:11-15):
    axis AX_X / AX_Y / AX_Z; a 2-D scene's cylinder has height 0 and lies in
    the z = 0 plane."""
    axis: int
    center: Tuple[float, float, float]
    radius: float
    height: float


@dataclass
class Material:
    name: str = "air"
    eps_r: float = 1.0
    mu_r: float = 1.0
    sigma_e: float = 0.0
    sigma_m: float = 0.0

    @staticmethod
    def air():
        return Material("XFDTD_AIR", 1, 1, 0, 0)

    @staticmethod
    def pec():
        # material/material.cpp:32-34
        return Material("XFDTD_PEC", 1, 1, 1e10, 0)


@dataclass
class DrudeMedium:
    """DrudeMedium(name, eps_inf, pole_omega, nv, sigma_e, sigma_m)
    (material/dispersive_material.h:112-113).  Its base Material is
    LinearDispersiveMaterial(name, 1, 1, sigma_e, sigma_m): eps_r = mu_r = 1
    in the standard coefficient arrays; the E field of its cells follows the
    Drude update with eps_inf, omega_p (rad/s) and the collision rate nu
    (1/s)."""
    name: str
    eps_inf: float
    omega_p: float
    nu: float
    sigma_e: float = 0.0
    sigma_m: float = 0.0

    @property
    def eps_r(self):
        return 1.0

    @property
    def mu_r(self):
        return 1.0


@dataclass
class DebyMedium:
    """DebyMedium(name, eps_s, eps_inf, tau, sigma_e, sigma_m)
    (material/dispersive_material.h:137-141): eps(w) = eps_inf +
    (eps_s - eps_inf) / (1 + i w tau); a LinearDispersiveMaterial
    (eps_r = mu_r = 1 in the standard arrays)."""
    name: str
    eps_s: float
    eps_inf: float
    tau: float
    sigma_e: float = 0.0
    sigma_m: float = 0.0

    @property
    def eps_r(self):
        return 1.0

    @property
    def mu_r(self):
        return 1.0


@dataclass
class LorentzMedium:
    """LorentzMedium(name, eps_s, eps_inf, omega_p, nv, sigma_e, sigma_m)
    (material/dispersive_material.h:59-76): eps(w) = eps_inf +
    (eps_s - eps_inf) omega_p^2 / (omega_p^2 + 2 i w nv - w^2); a
    LinearDispersiveMaterial (eps_r = mu_r = 1 in the standard arrays)."""
    name: str
    eps_s: float
    eps_inf: float
    omega_p: float
    nu: float
    sigma_e: float = 0.0
    sigma_m: float = 0.0

    @property
    def eps_r(self):
        return 1.0

    @property
    def mu_r(self):
        return 1.0


def dispersion_kind(m):
    """0 plain material, 1 Drude, 2 Debye, 3 Lorentz."""
    return {"DrudeMedium": 1, "DebyMedium": 2, "LorentzMedium": 3}.get(type(m).__name__, 0)


@dataclass
class Object:
    name: str
    shape: object
    material: object  # Material or DrudeMedium


@dataclass
class PML:
    orientation: int
    thickness: int
    order: int = 4
    sigma_ratio: float = 1.0
    alpha_min: float = 0.0
    alpha_max: float = 0.05
    kappa_max: float = 10.0


@dataclass
class GaussianWaveform:
    tau: float
    t_0: float
    amplitude: float = 1.0


@dataclass
class TFSF3D:
    distance_x: int
    distance_y: int
    distance_z: int
    e_0: float
    theta: float
    phi: float
    psi: float
    waveform: GaussianWaveform


@dataclass
class TFSF2D:
    """TFSF2D(distance_x, distance_y, phi_inc, ez_i0, waveform)
    (tfsf/tfsf_2d.cpp:10-24): a TMz plane wave travelling along phi_inc in
    the xy plane of a 2-D scene (theta = pi/2, psi = 0); its Ez is
    -ez_i0 * waveform (the polarisation vector of psi = 0 points along -z)."""
    distance_x: int
    distance_y: int
    phi_inc: float
    ez_i0: float
    waveform: object


@dataclass
class NffftFd:
    distance_x: int
    distance_y: int
    distance_z: int
    frequencies: Sequence[float]
    theta: Sequence[float]
    phi: Sequence[float]


@dataclass
class NffftBroadBand:
    """The NF-FF surface of NffftFd, observed in one direction (theta, phi)
    over every frequency bin of the run (2T-1 bins of 1/dt): the monostatic
    radar cross-section path."""
    distance_x: int
    distance_y: int
    distance_z: int
    theta: float
    phi: float


@dataclass
class SinusoidalWaveform:
    amplitude: float
    frequency: float
    phase: float = 0.0


@dataclass
class CosineModulatedGaussianWaveform:
    amplitude: float
    tau: float
    t_0: float
    modulation_frequency: float


# include/util/type_define.h Axis
AX_X, AX_Y, AX_Z = 0, 1, 2


@dataclass
class CurrentSource:
    cube: Cube
    orientation: int
    resistance: float
    waveform: object  # GaussianWaveform, SinusoidalWaveform or CosineModulatedGaussianWaveform


@dataclass
class VoltageSource:
    cube: Cube
    orientation: int
    resistance: float
    waveform: object


@dataclass
class Resistor:
    cube: Cube
    axis: int
    resistance: float


@dataclass
class Capacitor:
    cube: Cube
    axis: int
    capacitance: float


@dataclass
class Inductor:
    cube: Cube
    axis: int
    inductance: float


@dataclass
class ObjectPlane:
    """An Object whose cube has zero size along one axis: it sets sigma_e of
    the two tangential E components on that plane and nothing else."""
    name: str
    shape: Cube
    material: object


@dataclass
class VoltageMonitor:
    cube: Cube
    orientation: int
    output_dir: str = ""
    output_file_name: str = ""


@dataclass
class CurrentMonitor:
    cube: Cube
    orientation: int
    output_dir: str = ""
    output_file_name: str = ""


@dataclass
class Port:
    index: int
    impedance: complex
    is_source: bool
    voltage: VoltageMonitor
    current: CurrentMonitor


@dataclass
class Network:
    ports: List[Port]
    frequencies: Sequence[float]
    output_path: str = ""


def waveform_desc(w):
    """(kind, params) of a waveform: 0 Gaussian (amplitude, tau, t_0), 1
    sinusoid (amplitude, frequency, phase), 2 cosine-modulated Gaussian
    (amplitude, tau, t_0, frequency)."""
    if isinstance(w, SinusoidalWaveform):
        return 1, (w.amplitude, w.frequency, w.phase, 0.0)
    if isinstance(w, CosineModulatedGaussianWaveform):
        return 2, (w.amplitude, w.tau, w.t_0, w.modulation_frequency)
    return 0, (w.amplitude, w.tau, w.t_0, 0.0)


def lumped_desc(le):
    """(kind, dir, value, waveform kind, waveform params) of a lumped element:
    kind 0 CurrentSource, 1 VoltageSource, 2 Resistor, 3 Capacitor, 4 Inductor;
    waveform kind 0 Gaussian (amplitude, tau, t_0), 1 sinusoid (amplitude,
    frequency, phase), 2 cosine-modulated Gaussian (amplitude, tau, t_0,
    frequency)."""
    wave = waveform_desc
    if isinstance(le, CurrentSource):
        return (0, le.orientation, le.resistance) + wave(le.waveform)
    if isinstance(le, VoltageSource):
        return (1, le.orientation, le.resistance) + wave(le.waveform)
    if isinstance(le, Resistor):
        return (2, le.axis, le.resistance, 0, (0.0,) * 4)
    if isinstance(le, Capacitor):
        return (3, le.axis, le.capacitance, 0, (0.0,) * 4)
    if isinstance(le, Inductor):
        return (4, le.axis, le.inductance, 0, (0.0,) * 4)
    raise TypeError(f"not a lumped element: {le!r}")


@dataclass
class Probe:
    comp: int
    i: int
    j: int
    k: int


XY, YZ, ZX = range(3)  # PlaneType (util/type_define.h:30)


@dataclass
class TimeDomainFieldMonitor:
    """TimeDomainFieldMonitor(Cube, PlaneType, EMComponent, output_dir,
    output_file_name) (monitor/field_monitor.h:16-19): `comp` over the grid
    box of `cube`, one cell thick across `plane_type`, written as .npy at the
    end of the run."""
    cube: Cube
    plane_type: int
    comp: int
    output_dir: str
    output_file_name: str = ""


@dataclass
class MovieMonitor:
    """MovieMonitor(monitor, total_time_steps, plot_step)
    (monitor/movie_monitor.h): the wrapped monitor's slice as
    output_dir/%04d.npy after every step t with t % plot_step == 0 or
    t == total_time_steps - 1."""
    monitor: TimeDomainFieldMonitor
    total_time_steps: int
    plot_step: int


@dataclass
class Design:
    """Per-cell relative permittivity of a box of cells [i0,i1)x[j0,j1)x[k0,k1)
    (grid indices), lossless with mu_r = 1, applied after the objects, as one
    one-cell Object per cell would be (object/object.cpp:53-73).  eps_r is
    rounded to float like a Material's."""
    box: Tuple[int, int, int, int, int, int]
    eps_r: Sequence[float]  # C order over the box


@dataclass
class DftPlane:
    """Running DFT of one field component on a plane: axis fixed at `index`,
    the other two axes (in increasing order) over [lo1,hi1) x [lo2,hi2).
    After step t the sample is weighted by dt exp(-i 2 pi f t dt), the
    magnetic-current weight of nffft_fd.cpp:102-103.  The objective of the
    adjoint is J = sum |F|^2 over the plane."""
    comp: int
    axis: int
    index: int
    lo1: int
    hi1: int
    lo2: int
    hi2: int
    freq: float


@dataclass
class Scene:
    cell_size: float
    cfl: float = 0.99
    objects: List[Object] = field(default_factory=list)
    boundaries: List[PML] = field(default_factory=list)
    tfsf: Optional[TFSF3D] = None
    nffft: Optional[object] = None  # NffftFd or NffftBroadBand
    sources: List[object] = field(default_factory=list)  # lumped elements, in insertion order
    probes: List[Probe] = field(default_factory=list)
    monitors: List[object] = field(default_factory=list)  # field, movie, voltage and current monitors
    network: Optional[Network] = None
    design: Optional[Design] = None
    dft_plane: Optional[DftPlane] = None

    def add_object(self, o: Object):
        self.objects.append(o)
        return self

    def add_boundary(self, b: PML):
        self.boundaries.append(b)
        return self


def cpml_all_faces(thickness, order=3, sigma_ratio=0.8, alpha_min=0.0,
                   alpha_max=0.05, kappa_max=1.0):
    """Six CPML faces in the XN, XP, YN, YP, ZN, ZP insertion order."""
    return [PML(o, thickness, order, sigma_ratio, alpha_min, alpha_max, kappa_max)
            for o in range(6)]


def cube_domain_scene(n_total, dl, pml, cfl=0.9, objects=(), tfsf_gap=None,
                      nffft_gap=None, freqs=(), theta=(), phi=(),
                      incidence=(0.0, 0.0, 0.0), tau_steps=20.0):
    """A cubic vacuum domain of n_total cells including `pml` CPML cells per
    face, centred at the origin, like the reference's scatter examples
    (examples/3d_pec_sphere_bistatic_rcs.cpp:34-101).

    objects: extra Objects added after the vacuum domain.
    tfsf_gap: TFSF distance from the grid edge (None = no TFSF source).
    nffft_gap: NF-FF distance from the grid edge (None = no NF-FF surface).
    The Gaussian has tau = tau_steps*dt and t_0 = 4.5*tau.
    """
    n_in = n_total - 2 * pml
    half = n_in * dl / 2
    sc = Scene(cell_size=dl, cfl=cfl)
    sc.add_object(Object("domain", Cube((-half, -half, -half), (2 * half,) * 3),
                         Material.air()))
    for o in objects:
        sc.add_object(o)
    sc.boundaries = cpml_all_faces(pml)
    dt = courant_dt(dl, cfl)
    tau = tau_steps * dt
    if tfsf_gap is not None:
        th, ph, ps = incidence
        sc.tfsf = TFSF3D(tfsf_gap, tfsf_gap, tfsf_gap, 1.0, th, ph, ps,
                         GaussianWaveform(tau, 4.5 * tau, 1.0))
    if nffft_gap is not None:
        sc.nffft = NffftFd(nffft_gap, nffft_gap, nffft_gap, list(freqs),
                           list(theta), list(phi))
    return sc


def cylinder_2d_scene(nx, ny, pml, dl, radius, cfl=0.8, tfsf_gap=50, nffft_gap=40, phi_inc=PI / 4,
                      center_frequency=12e9, max_frequency=20e9, obs_phi=1.25 * PI, material=None):
    """The reference's 2-D scatter example (examples/2d_cylinder_scatter.cpp:
    26-106) at any size: an nx x ny vacuum plane with `pml` CPML cells on the
    four x and y faces, a z-axis Cylinder of `radius` (PEC unless `material`
    is given) at the plane's centre, a TFSF2D source along phi_inc with a
    cosine-modulated Gaussian (tau = 1.7 / (f_max - f_c), t_0 = 0.8 tau) and
    an NffftBroadBand surface looking along (pi/2, obs_phi)."""
    sc = Scene(cell_size=dl, cfl=cfl)
    sc.add_object(Object("free_space", Cube((0.0, 0.0, 0.0), (nx * dl, ny * dl, 0.0)), Material.air()))
    if radius > 0:
        mat = material if material is not None else Material("pec", 1, 1, 1e24, 0)
        sc.add_object(Object("objectA", Cylinder(AX_Z, (nx * 0.5 * dl, ny * 0.5 * dl, 0.0), radius, 0.0), mat))
    sc.boundaries = [PML(o, pml) for o in range(4)]
    tau = 1.7 / (max_frequency - center_frequency)
    if tfsf_gap is not None:
        sc.tfsf = TFSF2D(tfsf_gap, tfsf_gap, phi_inc, 1.0,
                         CosineModulatedGaussianWaveform(1.0, tau, 0.8 * tau, center_frequency))
    if nffft_gap is not None:
        sc.nffft = NffftBroadBand(nffft_gap, nffft_gap, 0, PI / 2, obs_phi)
    return sc


def courant_dt_2d(dl, cfl):
    """fdtd_basic_coff/fdtd_basic_coff.cpp:347-349 (a grid with nz = 1)."""
    import numpy as np
    c = float(np.float32(cfl))
    return c / (C_0 * math.sqrt(1.0 / (dl * dl) + 1.0 / (dl * dl)))


def courant_dt(dl, cfl):
    """fdtd_basic_coff/fdtd_basic_coff.cpp:342-346 with cfl rounded to float."""
    import numpy as np
    c = float(np.float32(cfl))
    return c / (C_0 * math.sqrt(3.0 / (dl * dl)))
