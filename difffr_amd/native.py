"""ctypes binding of libdfr_hip.so (the C ABI declared in include/dfr.h).

Loading fails loudly when the library is missing: there is no CPU fallback in
the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdfr_hip.so")

DFR_OK = 0
ERRORS = {-1: "DFR_ERR_ARG", -2: "DFR_ERR_HIP", -3: "DFR_ERR_CAPACITY", -4: "DFR_ERR_STATE", -5: "DFR_ERR_RANGE"}

# Every symbol include/dfr.h declares (checked by the CPU test suite).
EXPORTS = [
    "dfr_default_params", "dfr_create", "dfr_destroy", "dfr_last_error", "dfr_set_fluid", "dfr_add_boundary",
    "dfr_set_init_velocity", "dfr_init", "dfr_step", "dfr_reset", "dfr_get_rigid", "dfr_set_rigid",
    "dfr_get_jacobians", "dfr_get_net_gradients", "dfr_reset_gradient", "dfr_get_fluid", "dfr_get_field",
    "dfr_set_field", "dfr_get_boundary_field", "dfr_get_stats", "dfr_reset_stats", "dfr_neighbor_stats",
    "dfr_get_neighbors", "dfr_set_profiling", "dfr_kernel_times", "dfr_num_fluid", "dfr_num_boundary",
    "dfr_num_bodies", "dfr_adjoint_begin", "dfr_adjoint_backward", "dfr_adjoint_end", "dfr_adjoint_iterations",
    "dfr_set_fixed_iterations", "dfr_mlp_capture", "dfr_mlp_dataset_size", "dfr_mlp_dataset", "dfr_mlp_init_params",
    "dfr_mlp_set_params", "dfr_mlp_get_params", "dfr_mlp_train", "dfr_mlp_enable", "dfr_mlp_predict",
    "dfr_mlp_loss_grad", "dfr_mlp_loss_grad_noalloc", "dfr_mlp_input_grad", "dfr_set_halo_capacity", "dfr_halo_stats",
    "dfr_set_profiling_classes", "dfr_get_body", "dfr_set_gradient_mode", "dfr_count_1ring", "dfr_adjoint_start_gradient",
    "dfr_adjoint_segment_memory", "dfr_state_save", "dfr_state_load", "dfr_adjoint_arena_limit",
    "dfr_state_free", "dfr_adjoint_tape_bytes", "dfr_set_compute_gradient", "dfr_get_gradient_manager",
    "dfr_set_time",
]

MLP_NP = 2402


class Params(C.Structure):
    _fields_ = [
        ("radius", C.c_double), ("dt", C.c_double), ("rho0", C.c_double), ("gravity", C.c_double * 3),
        ("max_error", C.c_double), ("max_error_v", C.c_double),
        ("min_iterations", C.c_int32), ("max_iterations", C.c_int32), ("max_iterations_v", C.c_int32),
        ("use_pressure_warmstart", C.c_int32), ("use_divergence_warmstart", C.c_int32),
        ("enable_divergence_solver", C.c_int32), ("gradient_mode", C.c_int32),
        ("optimize_rotation", C.c_int32), ("gyroscopic", C.c_int32), ("compute_gradient", C.c_int32),
        ("uniform_acc_time", C.c_double), ("target_time", C.c_double),
        ("max_neighbors", C.c_int32), ("max_boundary_neighbors", C.c_int32), ("release_mode", C.c_int32),
        ("use_gradient_manager", C.c_int32),
        ("surface_tension_method", C.c_int32), ("surface_tension", C.c_double),
        ("surface_tension_boundary", C.c_double), ("viscosity_method", C.c_int32), ("viscosity", C.c_double),
        ("viscosity_boundary", C.c_double), ("cfl_method", C.c_int32), ("cfl_factor", C.c_double),
        ("cfl_min_time_step_size", C.c_double), ("cfl_max_time_step_size", C.c_double),
        ("use_rigid_contact", C.c_int32), ("rigid_contact_beta", C.c_double), ("rigid_contact_gamma", C.c_double),
        ("rigid_contact_support_radius_factor", C.c_double), ("rigid_contact_friction", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("iterations_v", C.c_int32), ("avg_density_err", C.c_double),
        ("avg_div_err", C.c_double), ("time", C.c_double), ("step_count", C.c_int64),
        ("total_iterations", C.c_int64), ("total_iterations_v", C.c_int64), ("steps_since_reset", C.c_int64),
        ("trajectory_finished", C.c_int32), ("dt", C.c_double),
    ]


_lib = None


def load(path: str | None = None):
    """Loads libdfr_hip.so and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libdfr_hip.so not found at {p}: build it with __graft_entry__.build() "
                           "(make -C difffr_amd/csrc); there is no CPU fallback")
    L = C.CDLL(p)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint32)
    L.dfr_default_params.argtypes = [C.POINTER(Params), C.c_double, C.c_double]
    L.dfr_default_params.restype = None
    L.dfr_create.argtypes = [C.POINTER(Params), C.c_int]
    L.dfr_create.restype = P
    L.dfr_destroy.argtypes = [P]
    L.dfr_last_error.argtypes = [P]
    L.dfr_last_error.restype = C.c_char_p
    L.dfr_set_fluid.argtypes = [P, C.c_int64, fp, fp]
    L.dfr_add_boundary.argtypes = [P, C.c_int64, fp, C.c_int, dp, dp, C.c_double]
    L.dfr_set_init_velocity.argtypes = [P, C.c_int, dp, dp]
    L.dfr_init.argtypes = [P]
    L.dfr_step.argtypes = [P, C.c_int]
    L.dfr_reset.argtypes = [P]
    for f in ("dfr_get_rigid", "dfr_get_jacobians", "dfr_get_net_gradients", "dfr_set_rigid"):
        getattr(L, f).argtypes = [P, C.c_int, dp]
    L.dfr_reset_gradient.argtypes = [P]
    L.dfr_set_time.argtypes = [P, C.c_double, C.c_int64]
    L.dfr_get_gradient_manager.argtypes = [P, C.c_int, C.c_int, dp, dp]
    L.dfr_get_fluid.argtypes = [P, fp, fp]
    L.dfr_get_field.argtypes = [P, C.c_char_p, fp]
    L.dfr_set_field.argtypes = [P, C.c_char_p, fp]
    L.dfr_get_boundary_field.argtypes = [P, C.c_int, C.c_char_p, fp]
    L.dfr_get_stats.argtypes = [P, C.POINTER(Stats)]
    L.dfr_reset_stats.argtypes = [P]
    L.dfr_neighbor_stats.argtypes = [P, dp]
    L.dfr_get_body.argtypes = [P, C.c_int, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dfr_set_gradient_mode.argtypes = [P, C.c_int]
    L.dfr_set_compute_gradient.argtypes = [P, C.c_int]
    L.dfr_count_1ring.argtypes = [P]
    L.dfr_count_1ring.restype = C.c_int64
    L.dfr_set_halo_capacity.argtypes = [P, C.c_int]
    L.dfr_set_profiling_classes.argtypes = [P, C.c_uint32]
    L.dfr_halo_stats.argtypes = [P, C.POINTER(C.c_int64)]
    L.dfr_get_neighbors.argtypes = [P, C.c_int, up, up, C.c_int64]
    L.dfr_get_neighbors.restype = C.c_int64
    L.dfr_set_profiling.argtypes = [P, C.c_int]
    L.dfr_kernel_times.argtypes = [P, C.c_int, C.c_void_p, dp, C.POINTER(C.c_int64)]
    L.dfr_num_fluid.argtypes = [P]
    L.dfr_num_fluid.restype = C.c_int64
    L.dfr_num_boundary.argtypes = [P, C.c_int]
    L.dfr_num_boundary.restype = C.c_int64
    L.dfr_num_bodies.argtypes = [P]
    ip = C.POINTER(C.c_int32)
    L.dfr_adjoint_begin.argtypes = [P, C.c_int]
    L.dfr_adjoint_backward.argtypes = [P, dp, dp]
    L.dfr_adjoint_end.argtypes = [P]
    L.dfr_adjoint_start_gradient.argtypes = [P, dp]
    L.dfr_adjoint_segment_memory.argtypes = [P, C.c_int64]
    L.dfr_adjoint_arena_limit.argtypes = [P, C.c_int64]
    L.dfr_adjoint_tape_bytes.argtypes = [P]
    L.dfr_adjoint_tape_bytes.restype = C.c_int64
    L.dfr_state_save.argtypes = [P, C.c_int]
    L.dfr_state_load.argtypes = [P, C.c_int]
    L.dfr_state_free.argtypes = [P, C.c_int]
    L.dfr_adjoint_iterations.argtypes = [P, ip, ip, C.c_int64]
    L.dfr_adjoint_iterations.restype = C.c_int64
    L.dfr_set_fixed_iterations.argtypes = [P, C.c_int, ip, ip]
    L.dfr_mlp_capture.argtypes = [P, C.c_int64, C.c_double, C.c_uint64]
    L.dfr_mlp_dataset_size.argtypes = [P]
    L.dfr_mlp_dataset_size.restype = C.c_int64
    L.dfr_mlp_dataset.argtypes = [P, fp, fp, C.c_int64]
    L.dfr_mlp_dataset.restype = C.c_int64
    L.dfr_mlp_init_params.argtypes = [P, C.c_uint64]
    L.dfr_mlp_set_params.argtypes = [P, fp, fp]
    L.dfr_mlp_get_params.argtypes = [P, fp, fp]
    L.dfr_mlp_train.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_uint64, dp]
    L.dfr_mlp_enable.argtypes = [P, C.c_int]
    L.dfr_mlp_predict.argtypes = [P, C.c_int64, fp, fp]
    L.dfr_mlp_loss_grad.argtypes = [P, C.c_int64, fp, fp, fp, dp]
    L.dfr_mlp_loss_grad_noalloc.argtypes = [P, C.c_int64, fp, fp, fp, dp]
    L.dfr_mlp_input_grad.argtypes = [P, C.c_int64, fp, fp, fp]
    _lib = L
    return L


def missing_exports(path: str | None = None):
    """Names from EXPORTS the shared object does not export (loads without a GPU)."""
    L = C.CDLL(path or LIB_PATH)
    out = []
    for name in EXPORTS:
        try:
            getattr(L, name)
        except AttributeError:
            out.append(name)
    return out
