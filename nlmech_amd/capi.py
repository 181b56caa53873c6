"""ctypes binding of the nlmech_amd C-ABI (include/nlmech_amd.h).

The structures below mirror the header field by field.  ``load()`` opens the
in-tree library nlmech_amd/lib/libnlmech_amd.so and raises if it is missing:
there is no CPU fallback behind this interface.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libnlmech_amd.so"

NLM_OK = 0
NLM_PREC_FP64_EXACT = 0
NLM_PREC_FP64 = 1
NLM_PREC_FP32_FORCE = 2
NLM_TIME_CENTRAL_DIFFERENCE = 0
NLM_TIME_VELOCITY_VERLET = 1
NLM_MATERIAL_RNP_BOND = 0
NLM_MATERIAL_ELASTIC_STATE = 1
NLM_QS_TANGENT_FULL = 0
NLM_QS_TANGENT_COMPONENT = 1
NLM_QS_TANGENT_FD = 2
NLM_QS_PRECOND_NONE = 0
NLM_QS_PRECOND_JACOBI = 1
NLM_QS_PRECOND_BLOCKS = 2
NLM_QS_PRECOND_COARSE = 16

REGION = {"rectangle": 0, "angled_rectangle": 1, "circle": 2, "torus": 3,
          "line": 4, "cuboid": 5}
TIME_FN = {"constant": 0, "linear": 1, "quadratic": 2, "sin": 3,
           "linear_step": 4, "linear_slow_fast": 5}
SPACE_FN = {"constant": 0, "sin_x": 1, "sin_y": 2, "sin_z": 3, "linear_x": 4,
            "linear_y": 5, "linear_z": 6, "hat_x": 7, "hat_y": 8, "hat_z": 9,
            "line_load": 10}


class MaterialDeck(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("Gc", C.c_double), ("KIc", C.c_double), ("plane_strain", C.c_int),
        ("compute_from_classical", C.c_int), ("n_bond_params", C.c_int),
        ("bond_params", C.c_double * 2), ("inf_type", C.c_int),
        ("n_inf_params", C.c_int), ("inf_params", C.c_double * 2),
        ("check_sc_factor", C.c_double), ("irreversible", C.c_int),
        ("no_penetration", C.c_int),
    ]


class RnpParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("C", C.c_double), ("beta", C.c_double),
        ("rbar", C.c_double), ("inv_factor", C.c_double),
        ("factor_sc", C.c_double), ("irreversible", C.c_int),
        ("inf_type", C.c_int), ("inf_a", C.c_double), ("inf_b", C.c_double),
        ("inf_moment", C.c_double), ("contact", C.c_int),
        ("contact_kn", C.c_double), ("contact_rc", C.c_double),
        ("E", C.c_double), ("K", C.c_double), ("nu", C.c_double),
        ("lambda_", C.c_double), ("G", C.c_double), ("Gc", C.c_double),
        ("KIc", C.c_double),
    ]


class EsDeck(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("G", C.c_double), ("plane_strain", C.c_int),
        ("critical_stretch", C.c_double),
    ]


class EsParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("G", C.c_double), ("nu", C.c_double), ("mu", C.c_double),
        ("factor2D", C.c_double), ("plane_strain", C.c_int),
        ("critical_stretch", C.c_double),
    ]


class FdProblem(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n", C.c_int64), ("x", C.POINTER(C.c_double)),
        ("vol", C.POINTER(C.c_double)), ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)), ("interior", C.POINTER(C.c_uint8)),
        ("dt", C.c_double), ("precision", C.c_int), ("time_scheme", C.c_int),
        ("mat", RnpParams), ("material", C.c_int), ("es", EsParams),
    ]


class Bc(C.Structure):
    _fields_ = [
        ("region", C.c_int), ("x1", C.c_double), ("x2", C.c_double),
        ("y1", C.c_double), ("y2", C.c_double), ("z1", C.c_double),
        ("z2", C.c_double), ("r1", C.c_double), ("r2", C.c_double),
        ("theta", C.c_double), ("ndir", C.c_int), ("dir", C.c_int * 3),
        ("time_fn", C.c_int), ("n_time_params", C.c_int),
        ("time_params", C.c_double * 4), ("space_fn", C.c_int),
        ("n_space_params", C.c_int), ("space_params", C.c_double * 4),
    ]


class Crack(C.Structure):
    _fields_ = [
        ("o", C.c_int), ("pb", C.c_double * 2), ("pt", C.c_double * 2),
        ("theta", C.c_double), ("activation_time", C.c_double),
    ]


class QsOptions(C.Structure):
    _fields_ = [
        ("nsteps", C.c_int), ("dt", C.c_double), ("tol_abs", C.c_double),
        ("tol_rel", C.c_double), ("max_newton", C.c_int),
        ("tangent", C.c_int), ("precond", C.c_int), ("lin_tol", C.c_double),
        ("lin_max_iter", C.c_int), ("predictor", C.c_int),
        ("gnn_refresh", C.c_int), ("perturbation", C.c_double),
    ]


class QsStats(C.Structure):
    _fields_ = [
        ("load_steps", C.c_int), ("newton_iters", C.c_int),
        ("linear_iters", C.c_int), ("linear_failures", C.c_int),
        ("residual", C.c_double), ("residual0", C.c_double),
        ("ms_total", C.c_double), ("ms_linear", C.c_double),
        ("ms_precond", C.c_double), ("ms_tangent", C.c_double),
    ]


# exported symbol -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U8 = C.POINTER(C.c_uint8)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
SIGNATURES = {
    "nlm_ctx_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "nlm_ctx_destroy": (None, [_P]),
    "nlm_last_error": (C.c_char_p, [_P]),
    "nlm_version": (C.c_char_p, []),
    "nlm_device_arch": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "nlm_rnp_params_from_deck": (C.c_int, [C.POINTER(MaterialDeck),
                                           C.POINTER(RnpParams)]),
    "nlm_neighbor_build": (C.c_int, [_P, C.c_int, C.c_int64, _D, C.c_double,
                                     _I64]),
    "nlm_neighbor_get": (C.c_int, [_P, _I64, _I32]),
    "nlm_fd_setup": (C.c_int, [_P, C.POINTER(FdProblem)]),
    "nlm_fd_set_ubc": (C.c_int, [_P, C.c_int, C.POINTER(Bc)]),
    "nlm_fd_set_fbc": (C.c_int, [_P, C.c_int, C.POINTER(Bc)]),
    "nlm_fd_add_cracks": (C.c_int, [_P, C.c_int, C.POINTER(Crack)]),
    "nlm_fd_set_state": (C.c_int, [_P, _D, _D, _D]),
    "nlm_fd_get_state": (C.c_int, [_P, _D, _D, _D]),
    "nlm_fd_set_bonds": (C.c_int, [_P, _U8]),
    "nlm_fd_get_bonds": (C.c_int, [_P, _U8]),
    "nlm_fd_bond_bytes_size": (C.c_int, [_P, _I64]),
    "nlm_fd_get_bond_bytes": (C.c_int, [_P, _U8]),
    "nlm_fd_get_fixity": (C.c_int, [_P, _U8]),
    "nlm_fd_get_time": (C.c_int, [_P, _D, _I64]),
    "nlm_fd_set_time": (C.c_int, [_P, C.c_double, C.c_int64]),
    "nlm_fd_init": (C.c_int, [_P]),
    "nlm_fd_step": (C.c_int, [_P, C.c_int]),
    "nlm_fd_compute_forces": (C.c_int, [_P]),
    "nlm_fd_postproc": (C.c_int, [_P, _F, _F, _D]),
    "nlm_fd_postproc_noalloc": (C.c_int, [_P, _F, _F, _D]),
    "nlm_fd_es_fields": (C.c_int, [_P, _F, _D, _D]),
    "nlm_fd_external_force": (C.c_int, [_P, C.c_double, _D]),
    "nlm_fd_reaction_force": (C.c_int, [_P, _D, C.c_double, _D, _D]),
    "nlm_fd_timed_steps": (C.c_int, [_P, C.c_int, _D, _D]),
    "nlm_fd_num_bonds": (C.c_int, [_P, _I64]),
    "nlm_fd_autotune": (C.c_int, [_P, C.POINTER(C.c_int), _D]),
    "nlm_fd_set_unroll": (C.c_int, [_P, C.c_int]),
    "nlm_fd_autotune_tiles": (C.c_int, [_P, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), _D]),
    "nlm_fd_set_tile": (C.c_int, [_P, C.c_int]),
    "nlm_fd_kernel_name": (C.c_char_p, [_P]),
    "nlm_es_params_from_deck": (C.c_int, [C.POINTER(EsDeck),
                                          C.POINTER(EsParams)]),
    "nlm_fd_get_node_fields": (C.c_int, [_P, _D, _D]),
    "nlm_fd_set_node_ubc": (C.c_int, [_P, C.c_int64, _I64, _U8, _D]),
    "nlm_fd_step_graph": (C.c_int, [_P, C.c_int, C.c_int]),
    "nlm_fd_timed_steps_graph": (C.c_int, [_P, C.c_int, C.c_int, _D, _D]),
    "nlm_qs_solve": (C.c_int, [_P, C.POINTER(QsOptions), C.POINTER(QsStats)]),
    "nlm_qs_assemble": (C.c_int, [_P, C.c_int]),
    "nlm_qs_get_tangent": (C.c_int, [_P, _D, _D]),
    "nlm_qs_apply": (C.c_int, [_P, _D, _D]),
    "nlm_qs_set_precond": (C.c_int, [_P, _D, _D]),
    "nlm_qs_get_precond": (C.c_int, [_P, _D, _D]),
    "nlm_qs_coarse_apply": (C.c_int, [_P, _D, _D]),
    "nlm_gnn_num_params": (C.c_int64, []),
    "nlm_gnn_create": (C.c_int, [_P, C.c_uint64]),
    "nlm_gnn_get_weights": (C.c_int, [_P, _F]),
    "nlm_gnn_set_weights": (C.c_int, [_P, _F]),
    "nlm_gnn_predict": (C.c_int, [_P]),
    "nlm_gnn_loss_grad": (C.c_int, [_P, C.c_int, _D, _D, _D, _F]),
    "nlm_gnn_adam_step": (C.c_int, [_P, C.c_double]),
    "nlm_gnn_zero_grad": (C.c_int, [_P]),
    "nlm_qs_linear_solve": (C.c_int, [_P, _D, _D, C.c_int, C.c_double,
                                      C.c_int, C.POINTER(C.c_int), _D]),
    "nlm_qs_set_graph": (C.c_int, [_P, C.c_int]),
    "nlm_qs_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
}

_lib = None


def load(path: Path | None = None) -> C.CDLL:
    """Open the HIP library; raises FileNotFoundError when it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} not found: build it with `python -m nlmech_amd.build` "
            "(the engine has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class NlmError(RuntimeError):
    pass
