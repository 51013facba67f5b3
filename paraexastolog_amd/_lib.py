"""ctypes binding of libesl.so (the C ABI declared in include/esl.h).

The library is loaded from this package directory only.  There is no fallback:
if libesl.so is missing or no HIP device is present, the calls fail with
EslError and the message says why.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libesl.so")

ESL_OK, ESL_E_IO, ESL_E_PARSE, ESL_E_ARG, ESL_E_DEVICE, ESL_E_NUMERIC, ESL_E_OOM, ESL_E_STATE = range(8)
STATUS_NAMES = {0: "OK", 1: "IO", 2: "PARSE", 3: "ARG", 4: "DEVICE", 5: "NUMERIC", 6: "OOM", 7: "STATE"}


class EslError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ESL_E_{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolveOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("tol", C.c_double),
        ("max_iters", C.c_int),
        ("restart", C.c_int),
        ("full", C.c_int),
        ("preconditioner", C.c_int),
        ("graph", C.c_int),
        ("warm_pi", C.POINTER(C.c_double)),
        ("inner_rtol", C.c_double),
        ("exact_fallback", C.c_int),
    ]


class SolveReport(C.Structure):
    _fields_ = [
        ("residual", C.c_double),
        ("iterations", C.c_int),
        ("refinements", C.c_int),
        ("reachable", C.c_int64),
        ("unknowns", C.c_int64),
        ("terminals_reached", C.c_int),
        ("sets_iterated", C.c_int),
        ("fp64_fallback", C.c_int),
        ("ms_total", C.c_double),
        ("ms_stage", C.c_double * 8),
        ("iterations_sum", C.c_int64),
        ("exact_solves", C.c_int),
    ]

    def as_dict(self):
        return {
            "residual": self.residual,
            "iterations": self.iterations,
            "refinements": self.refinements,
            "unknowns": self.unknowns,
            "terminals_reached": self.terminals_reached,
            "sets_iterated": self.sets_iterated,
            "fp64_fallback": self.fp64_fallback,
            "iterations_sum": self.iterations_sum,
            "exact_solves": self.exact_solves,
            "ms_total": self.ms_total,
            "ms_stage": list(self.ms_stage)[:5],
        }


# (name, restype, argtypes) of every entry point in include/esl.h and include/esl_mlp.h
_VP = C.c_void_p
_P = C.POINTER
SIGNATURES = [
    ("esl_version", C.c_char_p, []),
    ("esl_ctx_create", C.c_int, [C.c_int, _P(_VP)]),
    ("esl_ctx_destroy", None, [_VP]),
    ("esl_last_error", C.c_char_p, [_VP]),
    ("esl_model_load", C.c_int, [C.c_char_p, _P(_VP)]),
    ("esl_model_parse", C.c_int, [C.c_char_p, _P(_VP)]),
    ("esl_model_destroy", None, [_VP]),
    ("esl_model_num_nodes", C.c_int, [_VP]),
    ("esl_model_node_name", C.c_char_p, [_VP, C.c_int]),
    ("esl_model_rule", C.c_int, [_VP, C.c_int, _P(C.c_int), _P(C.c_int32), _P(C.c_uint8)]),
    ("esl_rates_fill", C.c_int, [_VP, C.c_int, C.c_double, C.c_double, _P(C.c_char_p), _P(C.c_double), C.c_int,
                                 _P(C.c_char_p), _P(C.c_double), C.c_int, _P(C.c_double)]),
    ("esl_initial_state", C.c_int, [_VP, _P(C.c_char_p), _P(C.c_uint8), C.c_int, C.c_double, _P(C.c_double)]),
    ("esl_table_build", C.c_int, [_VP, _VP, _P(C.c_double), _P(_VP)]),
    ("esl_graph_destroy", None, [_VP]),
    ("esl_graph_reorganize", C.c_int, [_VP, _VP]),
    ("esl_graph_import", C.c_int, [_VP, _VP, _P(C.c_int32), _P(C.c_int32), C.c_int64, C.c_int32]),
    ("esl_graph_query", C.c_int, [_VP, _P(C.c_int64)]),
    ("esl_graph_stage_counts", C.c_int, [_VP, _P(C.c_int64)]),
    ("esl_table_export", C.c_int, [_VP, _VP, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32)]),
    ("esl_graph_export", C.c_int, [_VP, _VP, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32)]),
    ("esl_solve_opts_default", None, [_P(SolveOpts)]),
    ("esl_solve", C.c_int, [_VP, _VP, _P(C.c_double), _P(C.c_double), _P(C.c_double), _P(SolveOpts),
                            _P(SolveReport)]),
    ("esl_solve_device", C.c_int, [_VP, _VP, _VP, _VP, _VP, _P(SolveOpts), _P(SolveReport)]),
    ("esl_solution_export", C.c_int, [_VP, _VP, _P(C.c_int32), _P(C.c_int32), _P(C.c_double), _P(C.c_int64),
                                      _P(C.c_int32), _P(C.c_int32), _P(C.c_double)]),
    ("esl_solve_batch_device", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _VP, _P(SolveOpts), _P(SolveReport)]),
    ("esl_device_alloc", C.c_int, [_VP, C.c_size_t, _P(_VP)]),
    ("esl_device_free", C.c_int, [_VP, _VP]),
    ("esl_memcpy_h2d", C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    ("esl_memcpy_d2h", C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    ("esl_memcpy_d2d", C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    ("esl_graph_stage_ms", C.c_int, [_VP, _P(C.c_double)]),
    ("esl_profile", C.c_int, [_VP, C.c_int]),
    ("esl_plan_create", C.c_int, [_VP, _VP, _VP, _P(_VP)]),
    ("esl_plan_destroy", None, [_VP]),
    ("esl_plan_query", C.c_int, [_VP, _P(C.c_int64)]),
    ("esl_plan_solve", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _P(SolveOpts), _P(SolveReport)]),
    ("esl_plan_solve_noalloc", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _P(SolveOpts), _P(SolveReport)]),
    ("esl_plan_gather_terminal", C.c_int, [_VP, _VP, C.c_int, _VP, _VP]),
    ("esl_plan_set_schedule", C.c_int, [_VP, C.c_int, C.c_int]),
    ("esl_plan_set_results", C.c_int, [_VP, C.c_int, _P(C.c_double), _P(C.c_int32), _P(C.c_int32)]),
    ("esl_plan_set_times", C.c_int, [_VP, _VP, C.c_int, _P(C.c_double), _P(C.c_double)]),
    ("esl_plan_autotune", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _P(SolveOpts), _P(C.c_int32), _P(C.c_double)]),
    ("esl_plan_solve_many", C.c_int, [_VP, C.c_int, _P(_VP), _P(C.c_int), _P(_VP), _P(_VP), _P(SolveOpts),
                                      _P(SolveReport)]),
    ("esl_mlp_create", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, _P(_VP)]),
    ("esl_mlp_destroy", None, [_VP]),
    ("esl_mlp_num_params", C.c_int64, [_VP]),
    ("esl_mlp_get_params", C.c_int, [_VP, _VP, _P(C.c_float)]),
    ("esl_mlp_set_params", C.c_int, [_VP, _VP, _P(C.c_float)]),
    ("esl_mlp_forward", C.c_int, [_VP, _VP, C.c_int, _VP, _VP]),
    ("esl_mlp_predict_warm", C.c_int, [_VP, _VP, C.c_int, _VP, _VP]),
    ("esl_mlp_gradients", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _P(C.c_float), _P(C.c_float)]),
    ("esl_mlp_train_step", C.c_int, [_VP, _VP, C.c_int, _VP, _VP, C.c_float, _P(C.c_float)]),
    ("esl_mlp_train_steps", C.c_int, [_VP, _VP, C.c_int, C.c_int, _VP, _VP, C.c_float, C.c_float, _P(C.c_float)]),
    ("esl_profile_stat", C.c_int, [_VP, C.c_int, C.c_char_p, C.c_int, _P(C.c_int64), _P(C.c_double),
                                   _P(C.c_double)]),
]

_lib = None


def lib():
    """Loads libesl.so once (raises EslError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EslError(ESL_E_DEVICE, f"{LIB_PATH} is not built; run python -m paraexastolog_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != ESL_OK:
        msg = lib().esl_last_error(ctx)
        raise EslError(status, msg.decode() if msg else "")
