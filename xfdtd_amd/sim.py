"""Python handle on the C++ Simulation through include/xfdtd_sim.h.

`Simulation(scene, total_steps, precision)` builds the reference's objects
from a scene description (xfdtd_amd/scene.py), runs the set-up on the CPU and
steps the fields on the GPU.  All numerical work happens in the HIP engine;
this module only moves arguments and result buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib

FAR_FIELD_NAMES = ["n_theta", "n_phi", "l_theta", "l_phi", "e_theta", "e_phi", "h_theta", "h_phi"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _index_of(seq, item):
    for q, x in enumerate(seq):
        if x is item:
            return q
    raise ValueError("a port's monitor must be one of the scene's monitors")


def _f64(v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64))


class Simulation:
    def __init__(self, scene, total_steps, precision=64, output_dir=None, plan_only=False):
        L = lib()
        self.scene = scene
        self.T = int(total_steps)
        self.precision = int(precision)
        self.h = L.xf_sim_create(float(scene.cell_size), float(scene.cfl))
        if not self.h:
            raise RuntimeError("xf_sim_create: " + L.xf_sim_last_error().decode())
        try:
            for o in scene.objects:
                m = o.material
                sh = o.shape
                dk = type(m).__name__
                if dk in ("DebyMedium", "LorentzMedium"):
                    if hasattr(sh, "height"):
                        shp, pp = 2, [*sh.center, sh.radius, sh.height, sh.axis]
                    elif hasattr(sh, "radius"):
                        shp, pp = 1, [*sh.center, sh.radius, 0.0, 0.0]
                    else:
                        shp, pp = 0, [*sh.origin, *sh.size]
                    mp = ([m.eps_s, m.eps_inf, m.tau, 0.0] if dk == "DebyMedium"
                          else [m.eps_s, m.eps_inf, m.omega_p, m.nu])
                    self._chk(L.xf_sim_add_dispersive_object(self.h, shp, _dp(_f64(pp)), 2 if dk == "DebyMedium" else 3,
                                                             _dp(_f64(mp)), float(m.sigma_e), float(m.sigma_m)))
                elif hasattr(m, "omega_p"):
                    drude = (float(m.eps_inf), float(m.omega_p), float(m.nu), float(m.sigma_e), float(m.sigma_m))
                    if hasattr(sh, "radius"):
                        self._chk(L.xf_sim_add_drude_sphere(self.h, _dp(_f64(sh.center)), float(sh.radius), *drude))
                    else:
                        self._chk(L.xf_sim_add_drude_cube(self.h, _dp(_f64(sh.origin)), _dp(_f64(sh.size)), *drude))
                elif type(o).__name__ == "ObjectPlane":
                    self._chk(L.xf_sim_add_plane(self.h, _dp(_f64(sh.origin)), _dp(_f64(sh.size)), m.eps_r,
                                                 m.mu_r, m.sigma_e, m.sigma_m))
                elif hasattr(sh, "height"):
                    self._chk(L.xf_sim_add_cylinder(self.h, int(sh.axis), _dp(_f64(sh.center)), float(sh.radius),
                                                    float(sh.height), m.eps_r, m.mu_r, m.sigma_e, m.sigma_m))
                elif hasattr(sh, "radius"):
                    c = _f64(sh.center)
                    self._chk(L.xf_sim_add_sphere(self.h, _dp(c), float(sh.radius), m.eps_r, m.mu_r,
                                                  m.sigma_e, m.sigma_m))
                else:
                    p, s = _f64(sh.origin), _f64(sh.size)
                    self._chk(L.xf_sim_add_cube(self.h, _dp(p), _dp(s), m.eps_r, m.mu_r, m.sigma_e,
                                                m.sigma_m))
            for b in scene.boundaries:
                self._chk(L.xf_sim_add_pml(self.h, int(b.orientation), int(b.thickness), int(b.order),
                                           b.sigma_ratio, b.alpha_min, b.alpha_max, b.kappa_max))
            t = scene.tfsf
            if t is not None and hasattr(t, "phi_inc"):
                from xfdtd_amd.scene import waveform_desc
                wk, wp = waveform_desc(t.waveform)
                self._chk(L.xf_sim_set_tfsf2d(self.h, int(t.distance_x), int(t.distance_y), float(t.phi_inc),
                                              float(t.ez_i0), wk, _dp(_f64(wp))))
            elif t is not None:
                dist = (C.c_size_t * 3)(t.distance_x, t.distance_y, t.distance_z)
                w = t.waveform
                self._chk(L.xf_sim_set_tfsf3d(self.h, dist, t.e_0, t.theta, t.phi, t.psi,
                                              w.amplitude, w.tau, w.t_0))
            n = scene.nffft
            if n is not None and not hasattr(n, "frequencies"):
                dist = (C.c_size_t * 3)(n.distance_x, n.distance_y, n.distance_z)
                od = output_dir.encode() if output_dir else None
                self._chk(L.xf_sim_set_nffft_broadband(self.h, dist, float(n.theta), float(n.phi), od))
            elif n is not None:
                dist = (C.c_size_t * 3)(n.distance_x, n.distance_y, n.distance_z)
                fr, th, ph = _f64(n.frequencies), _f64(n.theta), _f64(n.phi)
                od = output_dir.encode() if output_dir else None
                self._chk(L.xf_sim_set_nffft_fd(self.h, dist, _dp(fr), len(fr), _dp(th), len(th),
                                                _dp(ph), len(ph), od))
            from xfdtd_amd.scene import lumped_desc
            for s in scene.sources:
                p, sz = _f64(s.cube.origin), _f64(s.cube.size)
                kind, dr, val, wk, wp = lumped_desc(s)
                self._chk(L.xf_sim_add_lumped(self.h, kind, _dp(p), _dp(sz), int(dr), float(val), wk,
                                              _dp(_f64(wp))))
            for p in scene.probes:
                self._chk(L.xf_sim_add_probe(self.h, int(p.comp), p.i, p.j, p.k))
            self.line_monitors = []
            for m in getattr(scene, "monitors", []):
                lk = {"VoltageMonitor": 0, "CurrentMonitor": 1}.get(type(m).__name__)
                if lk is not None:
                    self._chk(L.xf_sim_add_line_monitor(self.h, lk, _dp(_f64(m.cube.origin)), _dp(_f64(m.cube.size)),
                                                        int(m.orientation), m.output_dir.encode(),
                                                        m.output_file_name.encode()))
                    self.line_monitors.append(m)
                    continue
                inner = m.monitor if hasattr(m, "plot_step") else m
                org, siz = _f64(inner.cube.origin), _f64(inner.cube.size)
                d = inner.output_dir.encode()
                if hasattr(m, "plot_step"):
                    self._chk(L.xf_sim_add_movie_monitor(self.h, _dp(org), _dp(siz), int(inner.plane_type),
                                                         int(inner.comp), d, int(m.total_time_steps),
                                                         int(m.plot_step)))
                else:
                    self._chk(L.xf_sim_add_field_monitor(self.h, _dp(org), _dp(siz), int(inner.plane_type),
                                                         int(inner.comp), d, inner.output_file_name.encode()))
            net = getattr(scene, "network", None)
            if net is not None:
                n = len(net.ports)
                ids = (C.c_int * n)(*[int(p.index) for p in net.ports])
                imp = _f64([x for p in net.ports for x in (complex(p.impedance).real, complex(p.impedance).imag)])
                src = (C.c_int * n)(*[1 if p.is_source else 0 for p in net.ports])
                vm = (C.c_int * n)(*[_index_of(self.line_monitors, p.voltage) for p in net.ports])
                im = (C.c_int * n)(*[_index_of(self.line_monitors, p.current) for p in net.ports])
                fr = _f64(net.frequencies)
                self._chk(L.xf_sim_set_network(self.h, n, ids, _dp(imp), src, vm, im, _dp(fr), len(fr),
                                               net.output_path.encode()))
            dz = getattr(scene, "design", None)
            if dz is not None:
                box = (C.c_size_t * 6)(*[int(v) for v in dz.box])
                er = np.ascontiguousarray(np.asarray(dz.eps_r, dtype=np.float32).ravel())
                self._chk(L.xf_sim_set_design(self.h, box, er.ctypes.data_as(C.POINTER(C.c_float))))
            dp_ = getattr(scene, "dft_plane", None)
            if dp_ is not None:
                self._chk(L.xf_sim_set_dft_plane(self.h, int(dp_.comp), int(dp_.axis), int(dp_.index),
                                                 int(dp_.lo1), int(dp_.hi1), int(dp_.lo2), int(dp_.hi2),
                                                 float(dp_.freq)))
            if plan_only:
                self._chk(L.xf_sim_plan(self.h, self.T))
            else:
                self._chk(L.xf_sim_init(self.h, self.T, self.precision))
        except Exception:
            L.xf_sim_destroy(self.h)
            self.h = None
            raise
        d = (C.c_size_t * 3)()
        L.xf_sim_dims(self.h, d)
        self.dims = (d[0], d[1], d[2])

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("xfdtd_amd: " + lib().xf_sim_last_error().decode())

    def _echk(self, rc):
        """Status of a direct engine call (xf_engine_*)."""
        if rc != 0:
            raise RuntimeError("xfdtd_amd engine: " + lib().xf_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().xf_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def dt(self):
        return lib().xf_sim_dt(self.h)

    @property
    def current_step(self):
        return lib().xf_sim_current_step(self.h)

    @property
    def engine(self):
        return lib().xf_sim_engine(self.h)

    def field_shape(self, comp):
        nx, ny, nz = self.dims
        return [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz),
                (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)][comp]

    def check_run(self):
        """Simulation::checkRun: write "simulation_check" (grid size) and
        "simulation_object_check" (the objects) in the working directory."""
        self._chk(lib().xf_sim_check_run(self.h, self.T))

    def plan_coeff(self, which):
        """Per-cell coefficient array `which` implied by the CPU plan."""
        out = np.empty(self.field_shape(which // 3), dtype=np.float64)
        self._chk(lib().xf_sim_plan_coeff(self.h, int(which), _dp(out)))
        return out

    def step(self, n):
        self._chk(lib().xf_sim_step(self.h, int(n)))

    def sync(self):
        self._chk(lib().xf_sim_sync(self.h))

    def get_field(self, comp):
        out = np.empty(self.field_shape(comp), dtype=np.float64)
        self._chk(lib().xf_sim_get_field(self.h, comp, _dp(out)))
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.field_shape(comp)
        self._chk(lib().xf_sim_set_field(self.h, comp, _dp(a)))

    def get_nffft(self):
        n = lib().xf_sim_nffft_size(self.h)
        out = np.empty(2 * n, dtype=np.float64)
        self._chk(lib().xf_sim_get_nffft(self.h, _dp(out)))
        return out.view(np.complex128)

    def farfield(self):
        n = self.scene.nffft
        shp = (len(n.frequencies), len(n.theta), len(n.phi))
        M = shp[0] * shp[1] * shp[2]
        out = np.empty(18 * M, dtype=np.float64)
        self._chk(lib().xf_sim_farfield(self.h, _dp(out)))
        cplx = out[:16 * M].view(np.complex128).reshape(8, *shp)
        pw = out[16 * M:].reshape(2, *shp)
        res = {k: cplx[q].copy() for q, k in enumerate(FAR_FIELD_NAMES)}
        res["power_theta"] = pw[0].copy()
        res["power_phi"] = pw[1].copy()
        return res

    # ---- broadband NF-FF (NffftBroadBand)
    def get_bb_samples(self):
        """Face-current series [T][8][Q] (xf_engine_get_broadband_samples)."""
        n = lib().xf_sim_broadband_samples_size(self.h)
        out = np.empty(n, dtype=np.float64)
        self._chk(lib().xf_sim_get_broadband_samples(self.h, _dp(out)))
        return out.reshape(self.T, 8, -1)

    def bb_farfield(self, k=None):
        """NffftBroadBand::calculateFarField over the 2T-1 bins: a dict of
        frequencies, wave_number, the surface integrals n_x .. l_z, the
        projections, e/h fields and power densities; k overrides the bins'
        wave numbers in the surface integrals."""
        Ns = 2 * self.T - 1
        out = np.empty(32 * Ns, dtype=np.float64)
        kk = None if k is None else _f64(k)
        self._chk(lib().xf_sim_broadband_farfield(self.h, None if kk is None else _dp(kk), _dp(out)))
        r = {"frequencies": out[:Ns].copy(), "wave_number": out[Ns:2 * Ns].copy()}
        cplx = out[2 * Ns:30 * Ns].view(np.complex128).reshape(14, Ns)
        for q, name in enumerate(["n_x", "n_y", "n_z", "l_x", "l_y", "l_z", "n_theta", "n_phi", "l_theta",
                                  "l_phi", "e_theta", "e_phi", "h_theta", "h_phi"]):
            r[name] = cplx[q].copy()
        r["power_theta"] = out[30 * Ns:31 * Ns].copy()
        r["power_phi"] = out[31 * Ns:].copy()
        if self.dims[2] == 1:
            ez = np.empty(2 * Ns, dtype=np.float64)
            self._chk(lib().xf_sim_broadband_ez(self.h, _dp(ez)))
            r["ez"] = ez.view(np.complex128).copy()
        return r

    # ---- voltage / current monitors and networks
    def get_line_monitors(self):
        """[monitor][step] values of the voltage and current monitors, in the
        order they appear among the scene's monitors."""
        out = np.empty((len(self.line_monitors), self.T), dtype=np.float64)
        if len(self.line_monitors):
            self._chk(lib().xf_sim_get_line_monitors(self.h, _dp(out)))
        return out

    def network_sparams(self):
        """{10 j + i: S_ji over the network's frequencies}."""
        L = lib()
        n = L.xf_sim_network_size(self.h)
        nf = len(self.scene.network.frequencies)
        keys = (C.c_int * max(n, 1))()
        out = np.empty(2 * max(n, 1) * nf, dtype=np.float64)
        self._chk(L.xf_sim_network_sparams(self.h, keys, _dp(out)))
        v = out.view(np.complex128).reshape(max(n, 1), nf)
        return {int(keys[q]): v[q].copy() for q in range(n)}

    def output_nffft(self):
        self._chk(lib().xf_sim_output_nffft(self.h))

    def output_data(self):
        """Simulation::outputData: NF-FF files (when the NF-FF has an output
        directory) and the field monitors' final slices."""
        self._chk(lib().xf_sim_output_nffft(self.h))

    def drude_box(self):
        b = (C.c_int64 * 6)()
        self._chk(lib().xf_sim_drude_box(self.h, b))
        return tuple(int(v) for v in b)

    def set_drude_j(self, comp, arr):
        """Set J of component comp over the Drude box."""
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(lib().xf_sim_set_drude_j(self.h, int(comp), _dp(a)))

    def get_drude_j(self, comp):
        """J of component comp over the Drude box [i0,i1)x[j0,j1)x[k0,k1)."""
        b = self.drude_box()
        out = np.empty((b[1] - b[0], b[3] - b[2], b[5] - b[4]), dtype=np.float64)
        self._chk(lib().xf_sim_get_drude_j(self.h, int(comp), _dp(out)))
        return out

    def get_probes(self):
        out = np.empty((len(self.scene.probes), self.T), dtype=np.float64)
        self._chk(lib().xf_sim_get_probes(self.h, _dp(out)))
        return out

    # ---- differentiation
    def get_dft(self):
        dp_ = self.scene.dft_plane
        out = np.empty(2 * (dp_.hi1 - dp_.lo1) * (dp_.hi2 - dp_.lo2), dtype=np.float64)
        self._chk(lib().xf_sim_get_dft(self.h, _dp(out)))
        return out.view(np.complex128).reshape(dp_.hi1 - dp_.lo1, dp_.hi2 - dp_.lo2)

    def grad(self, checkpoint_every=0):
        """(J, dJ/d eps_r over the design box): forward run from step 0 with
        checkpoints, then the adjoint (xf_engine_grad_eps; 0 = automatic
        checkpoint interval)."""
        b = self.scene.design.box
        g = np.empty((b[1] - b[0], b[3] - b[2], b[5] - b[4]), dtype=np.float64)
        J = C.c_double()
        self._chk(lib().xf_sim_grad_eps(self.h, int(checkpoint_every), _dp(g), C.byref(J)))
        return J.value, g

    def update_design(self, eps_r):
        """New eps_r (the design box's shape, any float array) for the design
        cells of an initialised simulation (xf_sim_update_design); the fields
        are left as they are."""
        b = self.scene.design.box
        e = np.ascontiguousarray(eps_r, dtype=np.float32)
        if e.size != (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]):
            raise ValueError(f"eps_r has {e.size} values for a design box of {(b[1] - b[0], b[3] - b[2], b[5] - b[4])}")
        self._chk(lib().xf_sim_update_design(self.h, e.ctypes.data_as(C.POINTER(C.c_float))))

    def adjoint_run(self, nsteps, fields):
        """Transpose of nsteps steps of the field map on the adjoint state: six
        fields, then (with Drude cells) the three J adjoints in E shapes."""
        packed = np.concatenate([np.ascontiguousarray(f, dtype=np.float64).ravel() for f in fields])
        self._chk(lib().xf_sim_adjoint_run(self.h, int(nsteps), _dp(packed), packed.size))
        out, off = [], 0
        for c in list(range(6)) + list(range(len(fields) - 6)):
            n = int(np.prod(self.field_shape(c)))
            out.append(packed[off:off + n].reshape(self.field_shape(c)).copy())
            off += n
        return out

    def checkpoint(self, slot):
        self._chk(lib().xf_sim_checkpoint(self.h, int(slot)))

    def restore(self, slot):
        self._chk(lib().xf_sim_restore(self.h, int(slot)))

    # ---- engine controls
    def set_timing(self, on=True):
        L = lib()
        if L.xf_engine_set_timing(self.engine, 1 if on else 0) != 0:
            raise RuntimeError(L.xf_last_error().decode())

    def kernel_times(self):
        """Average device time (ms) of the E and H update kernels per step."""
        L = lib()
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        if L.xf_engine_kernel_times(self.engine, C.byref(a), C.byref(b), C.byref(n)) != 0:
            raise RuntimeError(L.xf_last_error().decode())
        return a.value, b.value, n.value

    def set_sweep(self, mode=1, rows=0, planes=0):
        """Time-step scheme (xf_engine_set_sweep): 0 = split E and H kernels
        (the default), 1 = the fused E+H kernel where the scene allows it."""
        self._echk(lib().xf_engine_set_sweep(self.engine, int(mode), int(rows), int(planes)))

    @property
    def sweep_active(self):
        return bool(lib().xf_engine_sweep_active(self.engine))

    def set_divergence_check(self, every):
        """Read the device's non-finite flag every `every` steps (0: never);
        a diverged run makes step() raise, naming the step."""
        self._echk(lib().xf_engine_set_divergence_check(self.engine, int(every)))

    def set_adjoint_overlap(self, mode):
        """Gradient schedule (xf_engine_set_adjoint_overlap): -1 overlap the
        adjoint sweeps with the segment replays in FP32 only, 0 never, 1
        always."""
        self._echk(lib().xf_engine_set_adjoint_overlap(self.engine, int(mode)))

    @property
    def diverged_step(self):
        """The step at which non-finite H values were detected, or None."""
        s = lib().xf_engine_diverged_step(self.engine)
        return None if s < 0 else int(s)

    def set_tiling(self, tz, ty, tx, flags=0):
        L = lib()
        if L.xf_engine_set_tiling(self.engine, int(tz), int(ty), int(tx), int(flags)) != 0:
            raise RuntimeError(L.xf_last_error().decode())


class Batch:
    """Several set-up simulations of one precision stepped together on the
    GPU (include/xfdtd_engine.h, xf_batch_*): every per-step kernel is one
    launch over all of them.  Results equal stepping each one alone; read
    them through the member Simulations after `sync()`.  Close the batch
    before its members."""

    def __init__(self, sims):
        L = lib()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.engine for s in self.sims])
        h = C.c_void_p()
        if L.xf_batch_create(arr, len(self.sims), C.byref(h)) != 0:
            raise RuntimeError("xfdtd_amd: " + L.xf_last_error().decode())
        self.h = h

    def step(self, n):
        L = lib()
        if L.xf_batch_step(self.h, int(n)) != 0:
            raise RuntimeError("xfdtd_amd: " + L.xf_last_error().decode())

    def sync(self):
        L = lib()
        if L.xf_batch_sync(self.h) != 0:
            raise RuntimeError("xfdtd_amd: " + L.xf_last_error().decode())

    def grad(self, checkpoint_every=0):
        """[(J, dJ/d eps_r over the design box)] of every member, from step 0
        (xf_batch_grad_eps: the members' gradients in one batched run)."""
        L = lib()
        gs = []
        for s in self.sims:
            b = s.scene.design.box
            gs.append(np.empty((b[1] - b[0], b[3] - b[2], b[5] - b[4]), dtype=np.float64))
        ptrs = (C.POINTER(C.c_double) * len(gs))(*[_dp(g) for g in gs])
        J = np.empty(len(gs), dtype=np.float64)
        if L.xf_batch_grad_eps(self.h, int(checkpoint_every), ptrs, _dp(J)) != 0:
            raise RuntimeError("xfdtd_amd: " + L.xf_last_error().decode())
        return [(float(J[q]), gs[q]) for q in range(len(gs))]

    def close(self):
        if getattr(self, "h", None):
            lib().xf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
