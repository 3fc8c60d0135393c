"""ctypes loader of xfdtd_amd/libxfdtd_amd.so (the HIP engine + C++ host).

The library is built in-tree by `make -C xfdtd_amd/csrc` (or
__graft_entry__.build()).  There is no fallback: if it is missing, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxfdtd_amd.so")
_LIB = None

# every extern "C" symbol declared in include/xfdtd_engine.h and include/xfdtd_sim.h
ENGINE_SYMBOLS = [
    "xf_engine_create", "xf_engine_destroy", "xf_last_error", "xf_engine_set_materials",
    "xf_engine_set_dispersive", "xf_engine_set_cpml", "xf_engine_set_tfsf", "xf_engine_set_nffft", "xf_engine_nffft_size",
    "xf_engine_get_nffft", "xf_engine_farfield_sums", "xf_engine_farfield_sums_noalloc",
    "xf_engine_add_current_source",
    "xf_engine_add_probe", "xf_engine_get_probes", "xf_engine_step", "xf_engine_sync",
    "xf_engine_current_step", "xf_engine_get_field", "xf_engine_set_field",
    "xf_engine_kernel_times", "xf_engine_set_timing", "xf_engine_set_tiling", "xf_engine_set_sweep",
    "xf_engine_sweep_active", "xf_engine_set_divergence_check", "xf_engine_diverged_step",
    "xf_engine_set_adjoint_overlap",
    "xf_batch_create", "xf_batch_step", "xf_batch_sync", "xf_batch_size", "xf_batch_destroy", "xf_batch_grad_eps",
    "xf_engine_set_design", "xf_engine_update_design", "xf_engine_set_dft_plane", "xf_engine_get_dft", "xf_engine_checkpoint",
    "xf_engine_restore", "xf_engine_grad_eps", "xf_engine_adjoint_run", "xf_engine_set_drude",
    "xf_engine_drude_box", "xf_engine_get_drude_j", "xf_engine_set_horizon", "xf_engine_set_drude_j",
    "xf_engine_get_box", "xf_engine_set_nffft_broadband", "xf_engine_broadband_samples_size",
    "xf_engine_get_broadband_samples", "xf_engine_broadband_sums", "xf_engine_set_broadband_scratch",
    "xf_engine_add_inductor", "xf_engine_add_line_monitor", "xf_engine_line_monitor_count",
    "xf_engine_get_line_monitors",
]
SIM_SYMBOLS = [
    "xf_sim_last_error", "xf_sim_create", "xf_sim_destroy", "xf_sim_add_cube", "xf_sim_add_sphere",
    "xf_sim_add_pml", "xf_sim_set_tfsf3d", "xf_sim_add_dispersive_object", "xf_sim_add_cylinder", "xf_sim_set_tfsf2d", "xf_sim_set_nffft_fd", "xf_sim_add_current_source",
    "xf_sim_add_probe", "xf_sim_plan", "xf_sim_check_run", "xf_sim_plan_coeff", "xf_sim_init", "xf_sim_step", "xf_sim_sync", "xf_sim_current_step",
    "xf_sim_dims", "xf_sim_dt", "xf_sim_get_field", "xf_sim_set_field", "xf_sim_nffft_size",
    "xf_sim_get_nffft", "xf_sim_farfield", "xf_sim_output_nffft", "xf_sim_get_probes",
    "xf_sim_engine", "xf_sim_set_design", "xf_sim_update_design", "xf_sim_set_dft_plane", "xf_sim_get_dft", "xf_sim_grad_eps",
    "xf_sim_adjoint_run", "xf_sim_checkpoint", "xf_sim_restore", "xf_sim_add_drude_cube",
    "xf_sim_add_drude_sphere", "xf_sim_drude_box", "xf_sim_get_drude_j", "xf_sim_set_drude_j",
    "xf_sim_add_field_monitor", "xf_sim_add_movie_monitor", "xf_sim_set_nffft_broadband",
    "xf_sim_broadband_samples_size", "xf_sim_get_broadband_samples", "xf_sim_broadband_ez", "xf_sim_broadband_farfield",
    "xf_sim_add_lumped", "xf_sim_add_plane", "xf_sim_add_line_monitor", "xf_sim_get_line_monitors",
    "xf_sim_set_network", "xf_sim_network_size", "xf_sim_network_sparams",
]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `make -C xfdtd_amd/csrc` "
            "(or __graft_entry__.build()). There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    vp, d, sz, i, i64 = C.c_void_p, C.c_double, C.c_size_t, C.c_int, C.c_int64
    dp = C.POINTER(C.c_double)
    szp = C.POINTER(C.c_size_t)
    L.xf_last_error.restype = C.c_char_p
    L.xf_sim_last_error.restype = C.c_char_p
    L.xf_sim_create.restype = vp
    L.xf_sim_create.argtypes = [d, C.c_float]
    L.xf_sim_destroy.argtypes = [vp]
    L.xf_sim_add_cube.argtypes = [vp, dp, dp, C.c_float, C.c_float, d, d]
    L.xf_sim_add_sphere.argtypes = [vp, dp, d, C.c_float, C.c_float, d, d]
    L.xf_sim_add_pml.argtypes = [vp, i, i, i, d, d, d, d]
    L.xf_sim_set_tfsf3d.argtypes = [vp, szp, d, d, d, d, d, d, d]
    L.xf_sim_add_dispersive_object.argtypes = [vp, C.c_int, dp, C.c_int, dp, d, d]
    L.xf_sim_add_cylinder.argtypes = [vp, C.c_int, dp, d, d, C.c_float, C.c_float, d, d]
    L.xf_sim_set_tfsf2d.argtypes = [vp, C.c_size_t, C.c_size_t, d, d, C.c_int, dp]
    L.xf_sim_broadband_ez.argtypes = [vp, dp]
    L.xf_sim_set_nffft_fd.argtypes = [vp, szp, dp, i, dp, i, dp, i, C.c_char_p]
    L.xf_sim_add_current_source.argtypes = [vp, dp, dp, i, d, d, d, d]
    L.xf_sim_add_probe.argtypes = [vp, i, sz, sz, sz]
    L.xf_sim_init.argtypes = [vp, sz, i]
    L.xf_sim_plan.argtypes = [vp, sz]
    L.xf_sim_check_run.argtypes = [vp, sz]
    L.xf_sim_plan_coeff.argtypes = [vp, i, dp]
    L.xf_sim_step.argtypes = [vp, sz]
    L.xf_sim_sync.argtypes = [vp]
    L.xf_sim_current_step.restype = sz
    L.xf_sim_current_step.argtypes = [vp]
    L.xf_sim_dims.argtypes = [vp, szp]
    L.xf_sim_dt.restype = d
    L.xf_sim_dt.argtypes = [vp]
    L.xf_sim_get_field.argtypes = [vp, i, dp]
    L.xf_sim_set_field.argtypes = [vp, i, dp]
    L.xf_sim_nffft_size.restype = i64
    L.xf_sim_nffft_size.argtypes = [vp]
    L.xf_sim_get_nffft.argtypes = [vp, dp]
    L.xf_sim_farfield.argtypes = [vp, dp]
    L.xf_sim_output_nffft.argtypes = [vp]
    L.xf_sim_get_probes.argtypes = [vp, dp]
    L.xf_sim_engine.restype = vp
    L.xf_sim_engine.argtypes = [vp]
    L.xf_engine_kernel_times.argtypes = [vp, dp, dp, C.POINTER(C.c_int64)]
    L.xf_engine_set_timing.argtypes = [vp, i]
    L.xf_engine_set_tiling.argtypes = [vp, i, i, i, i]
    L.xf_engine_set_sweep.argtypes = [vp, i, i, i]
    L.xf_engine_sweep_active.argtypes = [vp]
    L.xf_engine_get_box.argtypes = [vp, i, C.POINTER(i64), dp]
    L.xf_engine_get_field.argtypes = [vp, i, dp]
    L.xf_engine_set_field.argtypes = [vp, i, dp]
    L.xf_engine_checkpoint.argtypes = [vp, i]
    L.xf_engine_restore.argtypes = [vp, i]
    L.xf_engine_set_divergence_check.argtypes = [vp, C.c_int]
    L.xf_engine_diverged_step.argtypes = [vp]
    L.xf_engine_set_adjoint_overlap.argtypes = [vp, C.c_int]
    L.xf_engine_diverged_step.restype = i64
    L.xf_engine_sync.argtypes = [vp]
    L.xf_engine_step.argtypes = [vp, i64]
    L.xf_engine_current_step.restype = i64
    L.xf_engine_current_step.argtypes = [vp]
    L.xf_batch_create.argtypes = [C.POINTER(vp), i, C.POINTER(vp)]
    L.xf_batch_step.argtypes = [vp, i64]
    L.xf_batch_sync.argtypes = [vp]
    L.xf_batch_size.argtypes = [vp]
    L.xf_batch_destroy.argtypes = [vp]
    L.xf_batch_grad_eps.argtypes = [vp, i, C.POINTER(dp), dp]
    L.xf_sim_set_design.argtypes = [vp, szp, C.POINTER(C.c_float)]
    L.xf_sim_update_design.argtypes = [vp, C.POINTER(C.c_float)]
    L.xf_sim_set_dft_plane.argtypes = [vp, i, i, sz, sz, sz, sz, sz, d]
    L.xf_sim_get_dft.argtypes = [vp, dp]
    L.xf_sim_grad_eps.argtypes = [vp, i, dp, dp]
    L.xf_sim_adjoint_run.argtypes = [vp, sz, dp, sz]
    L.xf_sim_checkpoint.argtypes = [vp, i]
    L.xf_sim_restore.argtypes = [vp, i]
    L.xf_sim_add_drude_cube.argtypes = [vp, dp, dp, d, d, d, d, d]
    L.xf_sim_add_drude_sphere.argtypes = [vp, dp, d, d, d, d, d, d]
    L.xf_sim_drude_box.argtypes = [vp, C.POINTER(C.c_int64)]
    L.xf_sim_get_drude_j.argtypes = [vp, i, dp]
    L.xf_sim_set_drude_j.argtypes = [vp, i, dp]
    L.xf_sim_add_field_monitor.argtypes = [vp, dp, dp, i, i, C.c_char_p, C.c_char_p]
    L.xf_sim_add_movie_monitor.argtypes = [vp, dp, dp, i, i, C.c_char_p, sz, sz]
    L.xf_sim_set_nffft_broadband.argtypes = [vp, szp, d, d, C.c_char_p]
    L.xf_engine_set_broadband_scratch.argtypes = [vp, i64]
    L.xf_engine_nffft_size.restype = i64
    L.xf_engine_nffft_size.argtypes = [vp]
    for name in ("xf_engine_farfield_sums", "xf_engine_farfield_sums_noalloc"):
        getattr(L, name).argtypes = [vp, d, dp, dp, i, dp, i, dp]
    ip = C.POINTER(C.c_int)
    L.xf_sim_add_lumped.argtypes = [vp, i, dp, dp, i, d, i, dp]
    L.xf_sim_add_plane.argtypes = [vp, dp, dp, C.c_float, C.c_float, d, d]
    L.xf_sim_add_line_monitor.argtypes = [vp, i, dp, dp, i, C.c_char_p, C.c_char_p]
    L.xf_sim_get_line_monitors.argtypes = [vp, dp]
    L.xf_sim_set_network.argtypes = [vp, i, ip, dp, ip, ip, ip, dp, i, C.c_char_p]
    L.xf_sim_network_size.argtypes = [vp]
    L.xf_sim_network_sparams.argtypes = [vp, ip, dp]
    L.xf_sim_broadband_samples_size.restype = i64
    L.xf_sim_broadband_samples_size.argtypes = [vp]
    L.xf_sim_get_broadband_samples.argtypes = [vp, dp]
    L.xf_sim_broadband_farfield.argtypes = [vp, dp, dp]
    _LIB = L
    return L
