"""ctypes binding of libuwg.so (the C ABI declared in include/uwg.h).

The library is loaded from the package directory.  There is no fallback: if the HIP
library is missing or no GPU is visible, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libuwg.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p

UWG_WARN_NOT_CONVERGED = 1
UWG_WARN_NONFINITE = 2


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precision", ctypes.c_int), ("learned", ctypes.c_int),
                ("outer_rtol", ctypes.c_double), ("inner_rtol", ctypes.c_double), ("mg_levels", ctypes.c_int),
                ("max_outer_its", ctypes.c_int), ("max_inner_its", ctypes.c_int), ("restart", ctypes.c_int),
                ("smooth_degree", ctypes.c_int), ("verbose", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("pressure_its", ctypes.c_int), ("velocity_presolve_its", ctypes.c_int),
                ("velocity_pressuresolve_its", ctypes.c_int), ("velocity_backsolve_its", ctypes.c_int),
                ("velocity_total_its", ctypes.c_int), ("outer_its", ctypes.c_int), ("vcycles", ctypes.c_int),
                ("levels", ctypes.c_int), ("converged", ctypes.c_int), ("fallback_used", ctypes.c_int),
                ("setup_time", ctypes.c_double), ("solve_time", ctypes.c_double), ("total_time", ctypes.c_double),
                ("res_momentum", ctypes.c_double), ("res_continuity", ctypes.c_double), ("norm_f", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, argtypes) for every exported entry point of include/uwg.h
SIGNATURES = {
    "uwg_last_error": [],
    "uwg_device_info": [ctypes.c_char_p, ctypes.c_int],
    "uwg_mesh_create": [ctypes.c_int, _ip, _dp, _dp, ctypes.POINTER(_vp)],
    "uwg_mesh_create_q2": [ctypes.c_int, _ip, _dp, _dp, ctypes.POINTER(_vp)],
    "uwg_set_q2_mfma": [ctypes.c_int],
    "uwg_set_q2_apply_kernel": [ctypes.c_int],
    "uwg_swarm_population_control": [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                     ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)],
    "uwg_advdiff_create": [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(_vp)],
    "uwg_advdiff_destroy": [_vp],
    "uwg_advdiff_set_coefficients": [_vp, _dp, _dp],
    "uwg_advdiff_set_scheme": [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int],
    "uwg_advdiff_integrate": [_vp, ctypes.c_double],
    "uwg_advdiff_max_dt": [_vp, _dp],
    "uwg_mesh_destroy": [_vp],
    "uwg_mesh_coords": [_vp, _dp],
    "uwg_mesh_min_separation": [_vp, _dp],
    "uwg_field_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    "uwg_field_destroy": [_vp],
    "uwg_field_set": [_vp, _dp],
    "uwg_field_get": [_vp, _dp],
    "uwg_swarm_create": [_vp, ctypes.POINTER(_vp)],
    "uwg_swarm_destroy": [_vp],
    "uwg_swarm_populate_random": [_vp, ctypes.c_int, ctypes.c_long],
    "uwg_swarm_populate_sobol": [_vp, ctypes.c_int, ctypes.c_long],
    "uwg_swarm_add_particles": [_vp, ctypes.c_long, _dp],
    "uwg_swarm_count": [_vp, _lp],
    "uwg_swarm_get_coords": [_vp, _dp],
    "uwg_swarm_set_coords": [_vp, _dp, _lp],
    "uwg_swarm_get_owners": [_vp, _ip],
    "uwg_swarm_get_cell_tables": [_vp, _ip, _ip],
    "uwg_swarm_set_material": [_vp, _ip],
    "uwg_swarm_update_owners": [_vp, _lp],
    "uwg_swarm_integration_points": [_vp, ctypes.c_int, ctypes.c_int],
    "uwg_swarm_get_integration": [_vp, _dp, _dp],
    "uwg_swarm_set_integration": [_vp, _dp, _dp],
    "uwg_advect_rk2": [_vp, _vp, ctypes.c_double, ctypes.c_int, _lp],
    "uwg_advect_max_dt": [_vp, _dp],
    "uwg_field_interpolate": [_vp, ctypes.c_long, _dp, _dp, _ip],
    "uwg_field_interpolate_noalloc": [_vp, ctypes.c_long, _dp, _dp, _ip],
    "uwg_stokes_create": [_vp, _vp, ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(_vp)],
    "uwg_stokes_destroy": [_vp],
    "uwg_stokes_set_materials": [_vp, ctypes.c_int, _dp, _dp],
    "uwg_stokes_set_particle_values": [_vp, ctypes.c_long, _dp, _dp],
    "uwg_swarm_set_var": [_vp, ctypes.c_int, ctypes.c_int, _dp],
    "uwg_stokes_get_particle_values": [_vp, ctypes.c_long, _dp, _dp],
    "uwg_stokes_eval_particle_functions": [_vp, _vp, ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_int,
                                           ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int],
    "uwg_stokes_assemble": [_vp, _vp],
    "uwg_stokes_assemble_fields": [_vp, _dp, _dp],
    "uwg_solve_opts_default": [ctypes.POINTER(SolveOpts)],
    "uwg_stokes_solve": [_vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
    "uwg_stokes_residuals": [_vp, _dp],
    "uwg_stokes_apply": [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "uwg_stokes_levels": [_vp, _ip, _ip],
    "uwg_stokes_get_assembly": [_vp, ctypes.c_int, _dp, _dp, _dp],
    "uwg_stokes_get_rhs": [_vp, _dp],
    "uwg_stokes_load_nn": [_vp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int],
    "uwg_stokes_set_nn_mode": [_vp, ctypes.c_int],
    "uwg_stokes_set_nn_monitor": [_vp, ctypes.c_int, ctypes.c_double],
    "uwg_stokes_set_nn_restart_rise": [_vp, ctypes.c_double],
    "uwg_stokes_set_velocity_block_solve": [_vp, ctypes.c_double],
    "uwg_stokes_capture": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "uwg_stokes_get_captured": [_vp, _fp, _ip],
    "uwg_stokes_level_solve": [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, _ip],
    "uwg_stokes_level_feature": [_vp, ctypes.c_int, _fp, _dp],
    "uwg_stokes_level_mask": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)],
    "uwg_nn_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                      ctypes.POINTER(_vp)],
    "uwg_nn_destroy": [_vp],
    "uwg_nn_param_count": [_vp, _lp],
    "uwg_nn_get_params": [_vp, _fp],
    "uwg_nn_set_params": [_vp, _fp],
    "uwg_nn_forward": [_vp, ctypes.c_int, _fp, _fp],
    "uwg_nn_loss_grad": [_vp, ctypes.c_int, _fp, _fp, _fp, _fp, _fp],
    "uwg_nn_set_data": [_vp, ctypes.c_long, _fp, _fp, _fp],
    "uwg_nn_train": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_long, _fp],
    "uwg_nn_eval": [_vp, ctypes.c_int, _fp],
    "uwg_snn_create": [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.POINTER(_vp)],
    "uwg_snn_destroy": [_vp],
    "uwg_snn_param_count": [_vp, ctypes.POINTER(ctypes.c_long)],
    "uwg_snn_get_params": [_vp, _fp],
    "uwg_snn_set_params": [_vp, _fp],
    "uwg_snn_features": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _fp],
    "uwg_snn_forward": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp],
    "uwg_snn_loss_grad": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _fp, _fp],
    "uwg_snn_add_samples": [_vp, _vp, ctypes.c_int, ctypes.c_long, ctypes.c_double],
    "uwg_snn_sample_count": [_vp, ctypes.POINTER(ctypes.c_long)],
    "uwg_snn_clear_data": [_vp],
    "uwg_snn_train": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_long, ctypes.c_long,
                      _fp],
    "uwg_snn_eval": [_vp, ctypes.c_long, ctypes.c_long, _fp],
    "uwg_stokes_load_schur_nn": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_long, _fp],
    "uwg_set_apply_chunk": [ctypes.c_int, ctypes.c_int],
    "uwg_set_cycle_bf16": [ctypes.c_int],
    "uwg_set_krylov_basis_fp32": [ctypes.c_int],
    "uwg_stokes_time_apply": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    "uwg_stream_copy": [ctypes.c_long, ctypes.c_int, _dp],
}


class UWGError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise UWGError(f"libuwg.so not found at {LIBPATH}; run __graft_entry__.build() "
                           f"(or python underworld2_amd/build.py)")
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.uwg_last_error.restype = ctypes.c_char_p
        L.uwg_solve_opts_default.restype = None
        _lib = L
    return _lib


def check(rc, what=""):
    """Map a C status to Python: <0 raises, >0 is returned for the caller to warn about."""
    if rc < 0:
        msg = lib().uwg_last_error().decode()
        raise UWGError(f"{what}: {msg}" if what else msg)
    return rc


def last_error():
    return lib().uwg_last_error().decode()


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def fptr(a):
    return a.ctypes.data_as(_fp)


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def device_info():
    buf = ctypes.create_string_buffer(256)
    check(lib().uwg_device_info(buf, 256), "device_info")
    return buf.value.decode()
