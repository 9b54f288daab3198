"""ctypes binding of libhns.so (the C ABI declared in include/hns.h)."""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HNS_LIBRARY", os.path.join(_PKG, "libhns.so"))

c_float_p = C.POINTER(C.c_float)
c_int32_p = C.POINTER(C.c_int32)


class hns_value_block(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("values", C.c_void_p)]


class hns_grid_data(C.Structure):
    _fields_ = [("size", C.c_size_t), ("coords", c_int32_p), ("num_blocks", C.c_size_t),
                ("blocks", C.POINTER(hns_value_block))]


class hns_leaf_grid(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("role", C.c_int32), ("num_leaves", C.c_size_t),
                ("origins", c_int32_p), ("values", c_float_p), ("active", C.POINTER(C.c_uint64))]


class hns_params(C.Structure):
    """CombustionParams (src/Cuda/Kernels.cuh:5-24), 72 bytes."""
    _fields_ = [("expansionRate", C.c_float), ("temperatureRelease", C.c_float), ("buoyancyStrength", C.c_float),
                ("ambientTemp", C.c_float), ("vorticityScale", C.c_float), ("factorScale", C.c_float),
                ("disturbanceEnable", C.c_float), ("disturbanceStrength", C.c_float),
                ("disturbanceSwirl", C.c_float), ("disturbanceThreshold", C.c_float),
                ("disturbanceGain", C.c_float), ("disturbanceFrequency", C.c_float), ("substeps", C.c_int32),
                ("gravity", C.c_float), ("maskDensityMin", C.c_float), ("maskDensityMax", C.c_float),
                ("downloadVel", C.c_int32), ("simReset", C.c_int32)]


class hns_solve_options(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("solver", C.c_int32), ("tol", C.c_float), ("max_iters", C.c_int32),
                ("check_every", C.c_int32)]


class hns_solve_stats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("iterations", C.c_int32), ("rel_residual", C.c_double),
                ("solve_ms", C.c_double), ("fallback_iterations", C.c_int32), ("reserved", C.c_int32)]


assert C.sizeof(hns_params) == 72

# (name, restype, argtypes) for every entry point of include/hns.h
_VP = C.c_void_p
_SIGS = [
    ("hns_last_error", C.c_char_p, []),
    ("hns_abi_version", C.c_int32, []),
    ("hns_create_index_grid", C.c_int32, [C.POINTER(hns_grid_data), C.c_float, C.POINTER(_VP)]),
    ("hns_topology_destroy", C.c_int32, [C.POINTER(_VP)]),
    ("hns_topology_info", C.c_int32, [_VP, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("hns_topology_neighbours", C.c_int32, [_VP, c_int32_p, C.c_size_t]),
    ("hns_leaf_order", C.c_int32, [c_int32_p, C.c_size_t, C.POINTER(C.c_uint64)]),
    ("hns_topology_builds", C.c_int32, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("hns_payload_build", C.c_int32, [C.POINTER(hns_leaf_grid), C.c_size_t, C.c_size_t, C.c_int32, C.POINTER(_VP)]),
    ("hns_payload_view", C.c_int32, [_VP, C.POINTER(hns_grid_data), c_int32_p]),
    ("hns_payload_destroy", C.c_int32, [C.POINTER(_VP)]),
    ("hns_payload_domain_mask", C.c_int32, [_VP, C.POINTER(C.c_uint64), C.c_size_t]),
    ("hns_write_leaf_grid", C.c_int32, [C.POINTER(hns_grid_data), C.c_char_p, C.POINTER(C.c_uint64), c_int32_p,
                                        c_float_p, C.POINTER(C.c_uint64)]),
    ("hns_scatter_leaves", C.c_int32, [C.POINTER(hns_grid_data), C.c_char_p, C.POINTER(C.c_uint8), C.c_float,
                                       c_int32_p, c_float_p, C.POINTER(C.c_uint64), C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    ("hns_ensure_context", C.c_int32, [C.POINTER(_VP), C.POINTER(hns_grid_data), _VP, C.c_float, C.c_int32]),
    ("hns_upload_sources", C.c_int32, [_VP, C.POINTER(hns_grid_data)]),
    ("hns_advance", C.c_int32, [_VP, C.c_int32, C.c_float, C.POINTER(hns_params)]),
    ("hns_advance_ex", C.c_int32, [_VP, C.c_int32, C.c_float, C.POINTER(hns_params), C.POINTER(hns_solve_options),
                                   C.POINTER(hns_solve_stats)]),
    ("hns_download_for_output", C.c_int32, [_VP, C.POINTER(hns_grid_data)]),
    ("hns_get_active_mask", C.c_int32, [_VP, C.POINTER(C.c_uint8), C.c_size_t, C.POINTER(C.c_size_t)]),
    ("hns_destroy_context", C.c_int32, [C.POINTER(_VP)]),
    ("hns_set_stream", C.c_int32, [_VP, _VP]),
    ("hns_context_info", C.c_int32, [_VP, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), c_int32_p]),
    ("hns_has_field", C.c_int32, [_VP, C.c_char_p]),
    ("hns_set_timing", C.c_int32, [_VP, C.c_int32]),
    ("hns_get_timing", C.c_int32, [_VP, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("hns_reset_timing", C.c_int32, [_VP]),
    ("hns_advect_index_grid", C.c_int32, [C.POINTER(hns_grid_data), C.c_float, C.c_float, _VP]),
    ("hns_advect_index_grid_velocity", C.c_int32, [C.POINTER(hns_grid_data), C.c_float, C.c_float, _VP]),
    ("hns_divergence", C.c_int32, [C.POINTER(hns_grid_data), C.c_float, _VP]),
    ("hns_project_non_divergent", C.c_int32, [C.POINTER(hns_grid_data), C.c_size_t, C.c_float, _VP]),
    ("hnsk_advect_velocity", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_float, _VP, _VP, C.c_int32]),
    ("hnsk_advect_scalars", C.c_int32, [_VP, _VP, C.c_int32, C.POINTER(_VP), C.POINTER(_VP), C.c_float, C.c_float,
                                        _VP, _VP, C.c_int32]),
    ("hnsk_advect_scalar", C.c_int32, [_VP, _VP, _VP, _VP, C.c_float, C.c_float, _VP, _VP, C.c_int32]),
    ("hnsk_vorticity_forces", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_float, C.c_float, C.c_float, _VP, C.c_float,
                                          _VP, C.c_int32]),
    ("hnsk_divergence", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_int32, _VP, C.c_int32]),
    ("hnsk_rbgs_sweep", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_int32, C.c_float, _VP, C.c_int32]),
    ("hnsk_residual", C.c_int32, [_VP, _VP, _VP, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double), _VP]),
    ("hnsk_residual_noalloc", C.c_int32, [_VP, _VP, _VP, C.c_float, _VP, _VP]),
    ("hnsk_pressure_solve", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_float, C.c_int32, C.POINTER(hns_solve_options),
                                        C.POINTER(hns_solve_stats), _VP]),
    ("hnsk_subtract_gradient", C.c_int32, [_VP, _VP, _VP, _VP, C.c_float, _VP, _VP, C.c_int32]),
    ("hnsk_enforce_collision", C.c_int32, [_VP, _VP, _VP, C.c_float, _VP, C.c_int32]),
    ("hnsk_combustion_buoyancy", C.c_int32, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_float, C.c_float,
                                             C.c_float, C.c_float, _VP, C.c_int32]),
    ("hnsk_to_colour_split", C.c_int32, [_VP, _VP, _VP, _VP, C.c_int32]),
    ("hnsk_from_colour_split", C.c_int32, [_VP, _VP, _VP, _VP, C.c_int32]),
    ("hnsk_aos_to_soa", C.c_int32, [C.c_size_t, _VP, _VP, _VP, C.c_int32]),
    ("hnsk_soa_to_aos", C.c_int32, [C.c_size_t, _VP, _VP, _VP, C.c_int32]),
    ("hns_precond_num_params", C.c_int64, []),
    ("hns_precond_create", C.c_int32, [_VP, C.c_int64, C.POINTER(_VP)]),
    ("hns_precond_create_device", C.c_int32, [_VP, C.c_int64, C.POINTER(_VP)]),
    ("hns_precond_refresh", C.c_int32, [_VP, _VP]),
    ("hns_precond_destroy", C.c_int32, [C.POINTER(_VP)]),
    ("hns_set_preconditioner", C.c_int32, [_VP, _VP]),
    ("hnsk_precond_apply", C.c_int32, [_VP, _VP, _VP, _VP, C.c_int32, _VP]),
    ("hnsk_precond_backward", C.c_int32, [_VP, _VP, _VP, _VP, _VP, _VP]),
    ("hnsk_precond_adam", C.c_int32, [C.c_int64, _VP, _VP, _VP, _VP, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_int32, _VP]),
    ("hnsk_poisson_apply", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_float, _VP, _VP, _VP]),
    ("hnsk_mg_apply", C.c_int32, [_VP, _VP, _VP, C.c_int32, _VP]),
    ("hnsk_pressure_solve_learned", C.c_int32, [_VP, _VP, _VP, _VP, C.c_float, C.POINTER(hns_solve_options),
                                                C.POINTER(hns_solve_stats), _VP]),
    ("hnsk_diff_forward", C.c_int32, [_VP, _VP, _VP, C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_float,
                                      C.c_int32, _VP, _VP, C.POINTER(_VP), _VP]),
    ("hnsk_diff_backward", C.c_int32, [_VP, _VP, _VP, _VP, _VP, _VP]),
    ("hns_diff_tape_info", C.c_int32, [_VP, c_int32_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("hns_diff_tape_destroy", C.c_int32, [C.POINTER(_VP)]),
    ("hnsk_advect_scalars_adj", C.c_int32, [_VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_float, _VP]),
    ("hnsk_advect_vector_adj", C.c_int32, [_VP, _VP, _VP, _VP, C.c_float, C.c_float, _VP]),
    ("hnsk_rbgs_adj", C.c_int32, [_VP, _VP, _VP, C.c_float, C.c_int32, C.c_float, _VP]),
]

EXPORTS = [s[0] for s in _SIGS]
_lib = None


def load() -> C.CDLL:
    """Loads libhns.so; raises ImportError with the build command when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libhns.so is not built (%s); run `python -m hnanosolver_amd.build`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
