"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        vp, d, sz, i = C.c_void_p, C.c_double, C.c_size_t, C.c_int
        dp = C.POINTER(C.c_double)
        szp = C.POINTER(C.c_size_t)
        L.xo_sim_create.restype = vp
        L.xo_sim_create.argtypes = [d, C.c_float]
        L.xo_sim_destroy.argtypes = [vp]
        L.xo_last_error.restype = C.c_char_p
        L.xo_sim_add_cube.argtypes = [vp, dp, dp, C.c_float, C.c_float, d, d]
        L.xo_sim_add_sphere.argtypes = [vp, dp, d, C.c_float, C.c_float, d, d]
        L.xo_sim_add_pml.argtypes = [vp, i, i, i, d, d, d, d]
        L.xo_sim_set_tfsf.argtypes = [vp, szp, d, d, d, d, d, d, d]
        L.xo_sim_set_nffft.argtypes = [vp, szp, dp, i, dp, i, dp, i]
        L.xo_sim_add_current_source.argtypes = [vp, dp, dp, i, d, d, d, d]
        L.xo_sim_add_probe.argtypes = [vp, i, sz, sz, sz]
        L.xo_sim_init.argtypes = [vp, sz]
        L.xo_sim_step.argtypes = [vp, sz]
        L.xo_sim_dims.argtypes = [vp, szp]
        L.xo_sim_dt.restype = d
        L.xo_sim_dt.argtypes = [vp]
        L.xo_sim_field_size.restype = sz
        L.xo_sim_field_size.argtypes = [vp, i]
        L.xo_sim_get_field.argtypes = [vp, i, dp]
        L.xo_sim_set_field.argtypes = [vp, i, dp]
        L.xo_sim_get_coeff.argtypes = [vp, i, dp]
        L.xo_sim_get_pml_profile.argtypes = [vp, i, dp]
        L.xo_sim_tfsf_aux_size.restype = sz
        L.xo_sim_tfsf_aux_size.argtypes = [vp]
        L.xo_sim_get_tfsf_aux.argtypes = [vp, dp, dp]
        L.xo_sim_get_waveform.argtypes = [vp, dp]
        L.xo_sim_nffft_size.restype = sz
        L.xo_sim_nffft_size.argtypes = [vp]
        L.xo_sim_get_nffft.argtypes = [vp, dp]
        L.xo_sim_farfield.argtypes = [vp, dp]
        L.xo_sim_get_probes.argtypes = [vp, dp]
        L.xo_sim_current_step.restype = sz
        L.xo_sim_current_step.argtypes = [vp]
        L.xo_set_threads.argtypes = [i]
        L.xo_set_threads.restype = i
        L.xo_sim_set_design.argtypes = [vp, szp, C.POINTER(C.c_float)]
        L.xo_sim_set_dft_plane.argtypes = [vp, i, i, sz, sz, sz, sz, sz, d]
        L.xo_sim_get_dft.argtypes = [vp, dp]
        L.xo_sim_grad.argtypes = [vp, dp, dp]
        L.xo_sim_adjoint_run.argtypes = [vp, sz, dp]
        L.xo_sim_set_drude.argtypes = [vp, d, d, d]
        L.xo_sim_get_drude_j.argtypes = [vp, i, dp]
        L.xo_sim_set_drude_j.argtypes = [vp, i, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr3(v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64))


def set_threads(n):
    return lib().xo_set_threads(int(n))


class OracleSim:
    """One FP64 oracle simulation built from a scene description (duck-typed:
    see xfdtd_amd/scene.py for the attribute names)."""

    def __init__(self, scene, total_steps):
        L = lib()
        self.scene = scene
        self.T = int(total_steps)
        self.h = L.xo_sim_create(float(scene.cell_size), float(scene.cfl))
        if not self.h:
            raise MemoryError("xo_sim_create")
        try:
            for o in scene.objects:
                m = o.material
                sh = o.shape
                if hasattr(sh, "radius"):
                    c = _arr3(sh.center)
                    self._chk(L.xo_sim_add_sphere(self.h, _dp(c), float(sh.radius),
                                                  m.eps_r, m.mu_r, m.sigma_e, m.sigma_m))
                else:
                    p, s = _arr3(sh.origin), _arr3(sh.size)
                    self._chk(L.xo_sim_add_cube(self.h, _dp(p), _dp(s), m.eps_r, m.mu_r,
                                                m.sigma_e, m.sigma_m))
                if hasattr(m, "omega_p"):
                    self._chk(L.xo_sim_set_drude(self.h, float(m.eps_inf), float(m.omega_p), float(m.nu)))
            for b in scene.boundaries:
                self._chk(L.xo_sim_add_pml(self.h, int(b.orientation), int(b.thickness),
                                           int(b.order), b.sigma_ratio, b.alpha_min,
                                           b.alpha_max, b.kappa_max))
            t = scene.tfsf
            if t is not None:
                dist = (C.c_size_t * 3)(t.distance_x, t.distance_y, t.distance_z)
                w = t.waveform
                self._chk(L.xo_sim_set_tfsf(self.h, dist, t.e_0, t.theta, t.phi, t.psi,
                                            w.amplitude, w.tau, w.t_0))
            n = scene.nffft
            if n is not None:
                dist = (C.c_size_t * 3)(n.distance_x, n.distance_y, n.distance_z)
                fr = _arr3(n.frequencies)
                th = _arr3(n.theta)
                ph = _arr3(n.phi)
                self._chk(L.xo_sim_set_nffft(self.h, dist, _dp(fr), len(fr), _dp(th), len(th),
                                             _dp(ph), len(ph)))
            for s in scene.sources:
                p, sz = _arr3(s.cube.origin), _arr3(s.cube.size)
                w = s.waveform
                self._chk(L.xo_sim_add_current_source(self.h, _dp(p), _dp(sz), int(s.orientation),
                                                      float(s.resistance), w.amplitude, w.tau,
                                                      w.t_0))
            for p in scene.probes:
                self._chk(L.xo_sim_add_probe(self.h, int(p.comp), p.i, p.j, p.k))
            dz = getattr(scene, "design", None)
            if dz is not None:
                box = (C.c_size_t * 6)(*[int(v) for v in dz.box])
                er = np.ascontiguousarray(np.asarray(dz.eps_r, dtype=np.float32).ravel())
                self._chk(L.xo_sim_set_design(self.h, box, er.ctypes.data_as(C.POINTER(C.c_float))))
            dp_ = getattr(scene, "dft_plane", None)
            if dp_ is not None:
                self._chk(L.xo_sim_set_dft_plane(self.h, int(dp_.comp), int(dp_.axis), int(dp_.index),
                                                 int(dp_.lo1), int(dp_.hi1), int(dp_.lo2), int(dp_.hi2),
                                                 float(dp_.freq)))
            self._chk(L.xo_sim_init(self.h, self.T))
        except Exception:
            L.xo_sim_destroy(self.h)
            self.h = None
            raise
        d = (C.c_size_t * 3)()
        L.xo_sim_dims(self.h, d)
        self.dims = (d[0], d[1], d[2])

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + lib().xo_last_error().decode())

    def close(self):
        if self.h:
            lib().xo_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def dt(self):
        return lib().xo_sim_dt(self.h)

    @property
    def current_step(self):
        return lib().xo_sim_current_step(self.h)

    def field_shape(self, comp):
        nx, ny, nz = self.dims
        return [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz),
                (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)][comp]

    def step(self, n):
        self._chk(lib().xo_sim_step(self.h, int(n)))

    def get_field(self, comp):
        out = np.empty(self.field_shape(comp), dtype=np.float64)
        self._chk(lib().xo_sim_get_field(self.h, comp, _dp(out)))
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.field_shape(comp)
        self._chk(lib().xo_sim_set_field(self.h, comp, _dp(a)))

    def get_coeff(self, which):
        out = np.empty(self.field_shape(which // 3), dtype=np.float64)
        self._chk(lib().xo_sim_get_coeff(self.h, which, _dp(out)))
        return out

    def get_pml_profile(self, b):
        d = self.scene.boundaries[b].thickness
        out = np.empty((6, d), dtype=np.float64)
        self._chk(lib().xo_sim_get_pml_profile(self.h, b, _dp(out)))
        return out

    def get_tfsf_aux(self):
        L = lib().xo_sim_tfsf_aux_size(self.h)
        e = np.empty(L, dtype=np.float64)
        h = np.empty(L - 1, dtype=np.float64)
        self._chk(lib().xo_sim_get_tfsf_aux(self.h, _dp(e), _dp(h)))
        return e, h

    def get_waveform(self):
        out = np.empty(self.T, dtype=np.float64)
        self._chk(lib().xo_sim_get_waveform(self.h, _dp(out)))
        return out

    def get_nffft(self):
        n = lib().xo_sim_nffft_size(self.h)
        out = np.empty(2 * n, dtype=np.float64)
        self._chk(lib().xo_sim_get_nffft(self.h, _dp(out)))
        return out.view(np.complex128)

    def farfield(self):
        n = self.scene.nffft
        M = len(n.frequencies) * len(n.theta) * len(n.phi)
        out = np.empty(18 * M, dtype=np.float64)
        self._chk(lib().xo_sim_farfield(self.h, _dp(out)))
        shp = (len(n.frequencies), len(n.theta), len(n.phi))
        cplx = out[:16 * M].view(np.complex128).reshape(8, *shp)
        pw = out[16 * M:].reshape(2, *shp)
        names = ["n_theta", "n_phi", "l_theta", "l_phi", "e_theta", "e_phi",
                 "h_theta", "h_phi"]
        res = {k: cplx[q].copy() for q, k in enumerate(names)}
        res["power_theta"] = pw[0].copy()
        res["power_phi"] = pw[1].copy()
        return res

    def get_dft(self):
        dp_ = self.scene.dft_plane
        out = np.empty(2 * (dp_.hi1 - dp_.lo1) * (dp_.hi2 - dp_.lo2), dtype=np.float64)
        self._chk(lib().xo_sim_get_dft(self.h, _dp(out)))
        return out.view(np.complex128).reshape(dp_.hi1 - dp_.lo1, dp_.hi2 - dp_.lo2)

    def grad(self):
        """(J, dJ/d eps_r over the design box) from a fresh simulation."""
        b = self.scene.design.box
        g = np.empty((b[1] - b[0], b[3] - b[2], b[5] - b[4]), dtype=np.float64)
        J = C.c_double()
        self._chk(lib().xo_sim_grad(self.h, _dp(g), C.byref(J)))
        return J.value, g

    def adjoint_run(self, nsteps, fields):
        """Reverse-run nsteps of the field map on the adjoint state: six
        fields, then (with Drude cells) the three J adjoints in E shapes."""
        packed = np.concatenate([np.ascontiguousarray(f, dtype=np.float64).ravel() for f in fields])
        self._chk(lib().xo_sim_adjoint_run(self.h, int(nsteps), _dp(packed)))
        out, off = [], 0
        for c in list(range(6)) + list(range(len(fields) - 6)):
            n = int(np.prod(self.field_shape(c)))
            out.append(packed[off:off + n].reshape(self.field_shape(c)).copy())
            off += n
        return out

    def set_drude_j(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(lib().xo_sim_set_drude_j(self.h, int(comp), _dp(a)))

    def get_drude_j(self, comp):
        out = np.empty(self.field_shape(comp), dtype=np.float64)
        self._chk(lib().xo_sim_get_drude_j(self.h, int(comp), _dp(out)))
        return out

    def get_probes(self):
        out = np.empty((len(self.scene.probes), self.T), dtype=np.float64)
        self._chk(lib().xo_sim_get_probes(self.h, _dp(out)))
        return out
