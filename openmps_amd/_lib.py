"""ctypes binding of the C ABI in include/mps.h (libmps_hip.so).

The library is loaded from this package directory only; if it is missing the
import fails loudly (there is no CPU fallback for the product path).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmps_hip.so")

_D = ctypes.c_double
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_I64 = ctypes.c_int64

MPS_OK = 0
MPS_ERR_CG_NOT_CONVERGED = 1
MPS_ERR_CELL_OVERFLOW = 2
MPS_ERR_HIP = 3
MPS_ERR_BAD_ARG = 4
MPS_ERR_NO_DEVICE = 5
MPS_ERR_STATE = 6

SOLVER_CG_FP64, SOLVER_MIXED_IR = 0, 1
STOP_INITIAL_RESIDUAL, STOP_RHS_NORM = 0, 1
WARM_PREVIOUS, WARM_ZERO, WARM_LEARNED = 0, 1, 2

AUX = {"n_without_spp": 0, "ecs": 1, "saved_x": 2, "saved_z": 3, "du_x": 4, "du_z": 5,
       "ppe_b": 6, "ppe_x": 7}

PRECOND_NONE, PRECOND_JACOBI, PRECOND_MULTILEVEL, PRECOND_AMG, PRECOND_AUTO = 0, 1, 2, 3, 4
ORDER_CELL, ORDER_KEEP = 0, 1

KERNELS = ("cell_keys", "cell_scan", "cell_scatter", "cell_sort", "neighbors", "density_ecs",
           "explicit", "update", "density_save", "ppe_count", "ppe_fill", "cg_init", "cg_spmv",
           "cg_update", "cg_iter32", "cg_init32", "ir_residual", "gradient", "ds", "other",
           "features", "mlp", "ml_setup", "ml_prec", "cg_spmv32", "cg_update32", "amg_cycle")


class MpsEnv(ctypes.Structure):
    _fields_ = [(n, _D) for n in (
        "max_dt_arg", "courant", "g", "rho", "nu", "r_e_by_l0", "l0", "min_x", "min_z", "max_x",
        "max_z", "r_lap_by_l0", "max_dt", "max_dx", "r_e", "r_lap", "gx", "gz", "neighbor_length",
        "n0", "n0_lap")]


class SolverOpts(ctypes.Structure):
    _fields_ = [("kind", _I), ("stop_rule", _I), ("warm_start", _I), ("precond", _I), ("tol", _D),
                ("max_iter", _I64), ("inner_tol", _D), ("inner_max_iter", _I64)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", _I64), ("outer_iterations", _I64), ("rr0", _D), ("rr", _D),
                ("rel_residual", _D), ("initial_rel_residual", _D), ("solve_ms", _D),
                ("learned_used", _I64), ("prev_rel_residual", _D),
                ("fallback", _I64), ("precond_used", _I64)]


class WallMotion(ctypes.Structure):
    _fields_ = [("vx", _D), ("vz", _D), ("ax", _D), ("az", _D), ("omega", _D), ("phase", _D)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", _I64), ("total_ms", _D),
                ("total_bytes", _D)]


WALL_FN = ctypes.CFUNCTYPE(None, _P, _D, _D, _SZ, ctypes.POINTER(ctypes.c_uint32),
                           ctypes.POINTER(_D), ctypes.POINTER(_D))
WALL_PRE_FN = ctypes.CFUNCTYPE(None, _P, _D, _D)

# name -> (restype, argtypes)
SIGNATURES = {
    "mps_env_init": (_I, [ctypes.POINTER(MpsEnv)] + [_D] * 11),
    "mps_env_set_laplacian_radius": (_I, [ctypes.POINTER(MpsEnv), _D]),
    "mps_create": (_I, [ctypes.POINTER(MpsEnv), _D, _I, ctypes.POINTER(_P)]),
    "mps_destroy": (None, [_P]),
    "mps_last_error": (ctypes.c_char_p, [_P]),
    "mps_get_env": (_I, [_P, ctypes.POINTER(MpsEnv)]),
    "mps_grid_shape": (_I, [_P, _P]),
    "mps_add_particles": (_I, [_P, _SZ] + [_P] * 7),
    "mps_size": (_SZ, [_P]),
    "mps_get_particles": (_I, [_P] + [_P] * 7),
    "mps_set_particles": (_I, [_P] + [_P] * 7),
    "mps_get_time": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "mps_set_time": (_I, [_P, _D, _D]),
    "mps_set_wall_callbacks": (_I, [_P, WALL_FN, WALL_PRE_FN, _P]),
    "mps_forward_time": (_I, [_P, _D]),
    "mps_forward_time_cfl": (_I, [_P, ctypes.POINTER(_D)]),
    "mps_determine_dt": (_I, [_P, ctypes.POINTER(_D)]),
    "mps_set_solver": (_I, [_P, ctypes.POINTER(SolverOpts)]),
    "mps_get_solver": (_I, [_P, ctypes.POINTER(SolverOpts)]),
    "mps_get_solve_stats": (_I, [_P, ctypes.POINTER(SolveStats)]),
    "mps_stage_search": (_I, [_P]),
    "mps_stage_density": (_I, [_P]),
    "mps_stage_error_correction": (_I, [_P]),
    "mps_stage_explicit": (_I, [_P]),
    "mps_stage_save_x": (_I, [_P]),
    "mps_stage_ppe_assemble": (_I, [_P]),
    "mps_stage_ppe_solve": (_I, [_P]),
    "mps_stage_pressure_writeback": (_I, [_P]),
    "mps_stage_pressure_gradient": (_I, [_P]),
    "mps_stage_dynamic_stabilize": (_I, [_P]),
    "mps_density_variation": (_I, [_P, _P]),
    "mps_get_neighbor_counts": (_I, [_P, _P]),
    "mps_get_neighbors": (_I, [_P, _P]),
    "mps_get_aux": (_I, [_P, _I, _P]),
    "mps_set_aux": (_I, [_P, _I, _P]),
    "mps_ppe_shape": (_I, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "mps_get_ppe": (_I, [_P, _P, _P, _P]),
    "mps_set_ppe": (_I, [_P, _SZ, _P, _P, _P, _P, _P]),
    "mps_set_profiling": (_I, [_P, _I]),
    "mps_get_kernel_stat": (_I, [_P, _I, ctypes.POINTER(KernelStat)]),
    "mps_reset_kernel_stats": (_I, [_P]),
    "mps_set_wall_motion": (_I, [_P, _SZ, ctypes.POINTER(WallMotion), ctypes.POINTER(ctypes.c_uint8)]),
    "mps_replay_spmv": (_I, [_P, _I, _I, ctypes.POINTER(_D)]),
    "mps_device_math_check": (_I, [_SZ, _P, _P, _P, _P]),
    "mps_device_copy_probe": (_I, [_SZ, ctypes.POINTER(_D)]),
    "mps_model_create": (_I, [_I, _I, ctypes.POINTER(_P)]),
    "mps_model_destroy": (None, [_P]),
    "mps_model_get_params": (_I, [_P] + [_P] * 6),
    "mps_model_set_params": (_I, [_P] + [_P] * 6),
    "mps_model_forward": (_I, [_P, _SZ, _P, _P]),
    "mps_model_gradients": (_I, [_P, _SZ, _P, _P, _P] + [_P] * 6),
    "mps_model_train": (_I, [_P, _P, _I, _I, _D, ctypes.c_uint64, _P]),
    "mps_dataset_create": (_I, [_I, _SZ, ctypes.POINTER(_P)]),
    "mps_dataset_destroy": (None, [_P]),
    "mps_dataset_size": (_I, [_P, ctypes.POINTER(_SZ)]),
    "mps_dataset_get": (_I, [_P, _SZ, _SZ, _P, _P]),
    "mps_attach_model": (_I, [_P, _P]),
    "mps_attach_dataset": (_I, [_P, _P, _D, ctypes.c_uint64]),
    "mps_get_features": (_I, [_P, _P]),
    "mps_set_storage_order": (_I, [_P, _I]),
    "mps_batch_create": (_I, [_I, _I, ctypes.POINTER(_P)]),
    "mps_batch_destroy": (None, [_P]),
    "mps_batch_last_error": (ctypes.c_char_p, [_P]),
    "mps_batch_add": (_I, [_P, _P]),
    "mps_batch_forward_time": (_I, [_P, _D, _P]),
    "mps_batch_solve_info": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "mps_cg_solve_csr_noalloc": (_I, [_I, _SZ, _P, _P, _P, _P, _P, _D, _I64, _P,
                                      ctypes.POINTER(_I64), ctypes.c_char_p, _SZ]),
}

_lib = None


def lib():
    """Loads libmps_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(python -m openmps_amd._build) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
