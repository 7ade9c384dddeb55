"""ctypes binding of libmetada_amd.so (include/metada_amd.h).

The structures below mirror the C declarations field by field.  Loading
fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmetada_amd.so")

MDA_OK = 0
MDA_E_ARG, MDA_E_DEVICE, MDA_E_NOMEM, MDA_E_NUMERIC, MDA_E_UNSUPPORTED = -1, -2, -3, -4, -5
ERRORS = {-1: "MDA_E_ARG", -2: "MDA_E_DEVICE", -3: "MDA_E_NOMEM", -4: "MDA_E_NUMERIC",
          -5: "MDA_E_UNSUPPORTED"}

COORD_GRID, COORD_GEOGRAPHIC, COORD_CARTESIAN = 0, 1, 2
HOP_CSR, HOP_HX = 0, 1
INFL_MULTIPLICATIVE, INFL_ADDITIVE, INFL_RELAXATION = 0, 1, 2
LW_LOC_GAUSSIAN, LW_LOC_EXPONENTIAL, LW_LOC_CUTOFF, LW_LOC_GASPARI_COHN = 0, 1, 2, 3
LW_GC_REFERENCE, LW_GC_STANDARD = 0, 1
LW_W_UNIFORM, LW_W_ADAPTIVE, LW_W_INVERSE_VAR, LW_W_LIKELIHOOD = 0, 1, 2, 3
LOC_GASPARI_COHN, LOC_CUTOFF = 0, 1
RW_REAL, RW_IDENTITY = 0, 1
INFLFORM_PRIOR, INFLFORM_POSTERIOR = 0, 1
SQRT_SYMMETRIC, SQRT_CHOLESKY = 0, 1
COMBINE_MATRIX, COMBINE_ELEMENTWISE = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
EIG_AUTO, EIG_JACOBI = 0, 1

_d, _i32, _i64 = C.c_double, C.c_int32, C.c_int64
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)


class Ensemble(C.Structure):
    _fields_ = [("k", _i32), ("nx", _i64), ("ny", _i64), ("nlev", _i64), ("data", C.c_void_p),
                ("members", C.c_void_p), ("on_device", _i32)]


class Obs(C.Structure):
    _fields_ = [("p", _i64), ("value", C.c_void_p), ("variance", C.c_void_p),
                ("loc_a", C.c_void_p), ("loc_b", C.c_void_p), ("loc_c", C.c_void_p),
                ("on_device", _i32)]


class Hop(C.Structure):
    _fields_ = [("kind", _i32), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("weight", C.c_void_p), ("nnz", _i64), ("hx", C.c_void_p), ("on_device", _i32)]


class Geometry(C.Structure):
    _fields_ = [("coord", _i32), ("col_a", C.c_void_p), ("col_b", C.c_void_p),
                ("col_c", C.c_void_p), ("on_device", _i32)]


class EnkfDiag(C.Structure):
    _fields_ = [("innovation_norm", _d), ("analysis_increment_norm", _d),
                ("background_spread", _d), ("analysis_spread", _d), ("max_kalman_gain", _d),
                ("min_kalman_gain", _d), ("condition_number", _d), ("ensemble_size", _i32),
                ("observation_count", _i64), ("inflation_method", _i32),
                ("inflation_factor", _d)]


class LwenkfParams(C.Structure):
    _fields_ = [("inflation", _d), ("localization_radius", _d), ("inflation_method", _i32),
                ("localization", _i32), ("gc_form", _i32), ("weighting", _i32)]


class LwenkfDiag(C.Structure):
    _fields_ = [("innovation_norm", _d), ("analysis_increment_norm", _d),
                ("background_spread", _d), ("analysis_spread", _d), ("max_kalman_gain", _d),
                ("min_kalman_gain", _d), ("condition_number", _d), ("localization_radius", _d),
                ("max_weight", _d), ("min_weight", _d), ("weight_variance", _d),
                ("ensemble_size", _i32), ("observation_count", _i64), ("inflation_method", _i32),
                ("localization", _i32), ("weighting", _i32), ("inflation_factor", _d),
                ("ms_gain_solve", _d), ("ms_update", _d)]


class LetkfParams(C.Structure):
    _fields_ = [("inflation", _d), ("radius", _d), ("localization", _i32), ("rweight", _i32),
                ("inflation_form", _i32), ("sqrt_form", _i32), ("combine", _i32),
                ("precision", _i32), ("rho_column", C.c_void_p), ("rho_on_device", _i32),
                ("eig_method", _i32), ("transform_precision", _i32), ("keep_transforms", _i32)]


class LetkfDiag(C.Structure):
    _fields_ = [("n_columns", _i64), ("n_empty_columns", _i64), ("total_local_obs", _i64),
                ("max_local_obs", _i64), ("max_jacobi_sweeps", _i32),
                ("n_fallback_columns", _i64), ("ms_obs", _d),
                ("ms_bucket", _d), ("ms_gram", _d), ("ms_solve", _d), ("ms_transform", _d),
                ("ms_total", _d), ("ms_eig_tridiag", _d), ("ms_eig_values", _d),
                ("ms_eig_vectors", _d), ("ms_solve_jacobi", _d)]


class BatchInfo(C.Structure):
    _fields_ = [("ms_total", _d), ("graph_replayed", _i32), ("n_graphs", _i32),
                ("gram_blocks", _i32 * 16), ("gram_queue", _i32 * 16)]


BATCH_GRAPH, BATCH_QUEUE, BATCH_AUTOTUNE = 1, 2, 4


class Metrics(C.Structure):
    _fields_ = [("rmse", _d), ("bias", _d), ("correlation", _d), ("crps", _d),
                ("avg_spread", _d)]


# every exported symbol with its signature; tests check that the library
# exports exactly what include/metada_amd.h declares
SIGNATURES = {
    "mda_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mda_ctx_destroy": (C.c_int, [C.c_void_p]),
    "mda_last_error": (C.c_char_p, [C.c_void_p]),
    "mda_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "mda_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "mda_ctx_workspace_bytes": (_i64, [C.c_void_p]),
    "mda_ensemble_mean_perturbations": (C.c_int, [C.c_void_p, C.POINTER(Ensemble), _d,
                                                  C.c_void_p, C.c_void_p, _i32]),
    "mda_obs_innovations": (C.c_int, [C.c_void_p, C.POINTER(Ensemble), C.POINTER(Obs),
                                      C.POINTER(Hop), C.c_void_p, C.c_void_p, C.c_void_p,
                                      _i32]),
    "mda_enkf_analyse": (C.c_int, [C.c_void_p, _d, _i32, C.POINTER(Ensemble), C.POINTER(Obs),
                                   C.POINTER(Hop), C.c_void_p, _i32, C.POINTER(EnkfDiag)]),
    "mda_lwenkf_analyse": (C.c_int, [C.c_void_p, C.POINTER(LwenkfParams), C.POINTER(Ensemble),
                                     C.POINTER(Obs), C.POINTER(Hop), C.c_void_p, _i32,
                                     C.c_void_p, C.POINTER(LwenkfDiag)]),
    "mda_etkf_analyse": (C.c_int, [C.c_void_p, _d, C.POINTER(Ensemble), C.POINTER(Obs),
                                   C.POINTER(Hop)]),
    "mda_etkf_analyse_noalloc": (C.c_int, [C.c_void_p, _d, C.POINTER(Ensemble), C.POINTER(Obs),
                                           C.POINTER(Hop)]),
    "mda_letkf_params_standard": (None, [C.POINTER(LetkfParams), _d, _d]),
    "mda_letkf_params_refcompat": (None, [C.POINTER(LetkfParams), _d, _d]),
    "mda_letkf_analyse": (C.c_int, [C.c_void_p, C.POINTER(LetkfParams), C.POINTER(Ensemble),
                                    C.POINTER(Geometry), C.POINTER(Obs), C.POINTER(Hop),
                                    C.POINTER(LetkfDiag)]),
    "mda_letkf_analyse_batch": (C.c_int, [C.c_void_p, _i32, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, _i32,
                                          C.POINTER(BatchInfo)]),
    "mda_letkf_last_transforms": (C.c_int, [C.c_void_p, C.c_void_p, _i64, C.c_void_p]),
    "mda_letkf_last_local_counts": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mda_letkf_last_column_flags": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mda_ensemble_metrics": (C.c_int, [C.c_void_p, C.POINTER(Ensemble), C.c_void_p, _i32,
                                       C.POINTER(Metrics)]),
    "mda_l96_forecast": (C.c_int, [C.c_void_p, C.c_void_p, _i64, _i64, _d, _d, _i32, _i32,
                                   _i32]),
    "mda_letkf_inflation_features": (C.c_int, [C.c_void_p, C.POINTER(LetkfParams),
                                               C.POINTER(Ensemble), C.POINTER(Geometry),
                                               C.POINTER(Obs), C.POINTER(Hop), C.c_void_p,
                                               C.c_void_p, _i32]),
    "mda_last_io_error": (C.c_char_p, []),
    "mda_hop_idw4_build": (C.c_int, [C.c_void_p, C.POINTER(Obs), _i32, C.POINTER(Geometry), _i64,
                                     _i64, _i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     _i32]),
    "mda_obs_read_grid_file": (C.c_int, [C.c_void_p, C.c_char_p, _d, _d, _i32,
                                         C.POINTER(C.c_void_p)]),
    "mda_obs_set_view": (C.c_int, [C.c_void_p, C.POINTER(Obs), C.POINTER(_i32)]),
    "mda_obs_set_free": (None, [C.c_void_p]),
    "mda_state_read_text": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, _i64,
                                      C.POINTER(_i64)]),
    "mda_state_write_text": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, _i64, _i64]),
    "mda_state_write_binary": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, _i64, _i64, _i64]),
    "mda_state_read_binary": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, _i64,
                                        C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "mda_ensemble_save": (C.c_int, [C.c_void_p, C.POINTER(Ensemble), C.c_char_p, C.c_char_p,
                                    _i32]),
    "mda_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _i64, C.c_void_p, _i32]),
    "mda_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _i64,
                                    C.POINTER(_d), C.c_void_p, _i32]),
    "mda_mlp_adam_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    _i64, C.c_float, C.c_float, C.c_float, C.c_float, _i32]),
}
MLP_NIN, MLP_HIDDEN, MLP_NPARAM = 8, 64, 4801

_lib = None


class MdaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load():
    """Loads the built library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime with the same SONAME
        # (libamdhip64.so.7); loading torch first makes this library bind to
        # that runtime, so one process never holds two HIP runtimes.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m metada_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(ctx, rc):
    if rc != MDA_OK:
        msg = load().mda_last_error(ctx)
        raise MdaError(rc, msg.decode() if msg else "")
    return rc
