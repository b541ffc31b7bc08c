"""ctypes binding of libpdfp_hip (include/pdfp_hip.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing, importing the solver raises, naming the build
command.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PDFP_HIP_LIB selects another build of the library (e.g. a variant built
# with `python -m pdfp_amd.build --out PATH -D NAME=VALUE`).
LIB_PATH = os.environ.get("PDFP_HIP_LIB", os.path.join(HERE, "libpdfp_hip.so"))
PROBE_PATH = os.path.join(HERE, "libpdfp_probe.so")

TERMINATION = {0: "UNSPECIFIED", 1: "OPTIMAL", 2: "ITERATION_LIMIT", 3: "TIME_LIMIT",
               4: "PRIMAL_INFEASIBLE", 5: "DUAL_INFEASIBLE"}
PRECISION = {"fp64": 0, "fp32": 1}


class PdfpParams(C.Structure):
    _fields_ = [
        ("eps_optimal_relative", C.c_double),
        ("eps_feasible_relative", C.c_double),
        ("iteration_limit", C.c_int64),
        ("time_limit", C.c_double),
        ("termination_evaluation_frequency", C.c_int32),
        ("sv_max_iterations", C.c_int32),
        ("reflection_coefficient", C.c_double),
        ("sv_tolerance", C.c_double),
        ("artificial_restart_threshold", C.c_double),
        ("sufficient_reduction_for_restart", C.c_double),
        ("necessary_reduction_for_restart", C.c_double),
        ("k_p", C.c_double),
        ("k_i", C.c_double),
        ("k_d", C.c_double),
        ("i_smooth", C.c_double),
    ]


class PdfpInfo(C.Structure):
    _fields_ = [
        ("relative_primal_residual", C.c_double),
        ("relative_dual_residual", C.c_double),
        ("relative_l1_gap", C.c_double),
        ("l1_distance", C.c_double),
        ("fixed_point_error", C.c_double),
        ("initial_fixed_point_error", C.c_double),
        ("primal_weight", C.c_double),
        ("step_size", C.c_double),
        ("max_singular_value", C.c_double),
        ("solve_seconds", C.c_double),
        ("total_count", C.c_int64),
        ("termination_reason", C.c_int32),
        ("n_restarts", C.c_int32),
        ("precision", C.c_int32),
        ("n_evaluations", C.c_int32),
        ("npb_a", C.c_int32),
        ("npb_at", C.c_int32),
        ("persist_a", C.c_int32),
        ("persist_at", C.c_int32),
        ("small_path", C.c_int32),
        ("lane_rows_a", C.c_int32),
        ("lane_rows_at", C.c_int32),
        ("block_graph", C.c_int32),
        ("column_slices", C.c_int32),
        ("slice_min_len", C.c_int32),
        ("nontemporal", C.c_int32),
    ]


class PdlpInfo(C.Structure):
    _fields_ = [
        ("relative_primal_residual", C.c_double),
        ("relative_dual_residual", C.c_double),
        ("relative_optimality_gap", C.c_double),
        ("objective_value", C.c_double),
        ("fixed_point_error", C.c_double),
        ("primal_weight", C.c_double),
        ("step_size", C.c_double),
        ("solve_seconds", C.c_double),
        ("total_count", C.c_int64),
        ("termination_reason", C.c_int32),
        ("n_restarts", C.c_int32),
        ("small_path", C.c_int32),
    ]


# Every C-ABI entry point: name -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
SIGNATURES = {
    "pdfp_default_params": (None, [C.POINTER(PdfpParams)]),
    "pdfp_last_error": (C.c_char_p, [_P]),
    "pdfp_create": (C.c_int, [C.POINTER(_P), C.c_int64, C.c_int64, C.c_int64, _I64, _I64, _D,
                              C.c_int32, _D, _U8, _U8, _D, _D, _D, C.c_int32, C.c_int32]),
    "pdfp_set_original_primal_residual": (C.c_int, [_P, _D]),
    "pdfp_set_x": (C.c_int, [_P, _D]),
    "pdfp_set_dual": (C.c_int, [_P, _D, _D]),
    "pdfp_set_target": (C.c_int, [_P, _D]),
    "pdfp_set_column_scale": (C.c_int, [_P, _D]),
    "pdfp_set_target_original": (C.c_int, [_P, _D]),
    "pdfp_warm_start": (C.c_int, [_P]),
    "pdfp_solve": (C.c_int, [_P, C.POINTER(PdfpParams), C.c_double, C.POINTER(C.c_int32), _I64]),
    "pdfp_get_x": (C.c_int, [_P, _D]),
    "pdfp_get_vector": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_get_info": (C.c_int, [_P, C.POINTER(PdfpInfo)]),
    "pdfp_destroy": (None, [_P]),
    "pdlp_last_error": (C.c_char_p, [_P]),
    "pdlp_create": (C.c_int, [C.POINTER(_P), C.c_int64, C.c_int64, C.c_int64, _I64, _I64, _D,
                              C.c_int32, _D, _D, _D, _D, _D, C.c_double, C.c_double, C.c_int32]),
    "pdlp_set_cupdlpx_termination": (C.c_int, [_P, _D, _D]),
    "pdlp_solve": (C.c_int, [_P, C.POINTER(PdfpParams), C.POINTER(C.c_int32), _I64]),
    "pdlp_set_kernel_path": (C.c_int, [_P, C.c_int32]),
    "pdlp_get_primal": (C.c_int, [_P, _D]),
    "pdlp_get_dual_slack": (C.c_int, [_P, _D]),
    "pdlp_get_info": (C.c_int, [_P, C.POINTER(PdlpInfo)]),
    "pdlp_destroy": (None, [_P]),
    "pdfp_max_singular_value": (C.c_int, [_P, C.c_int32, C.c_double, _D]),
    "pdfp_start_vector": (None, [C.c_int64, C.c_uint64, _D]),
    "pdfp_spmv": (C.c_int, [_P, C.c_int32, _D, _D]),
    "pdfp_time_iterations": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_time_kernels": (C.c_int, [_P, C.c_int32, _D, _D]),
    "pdfp_time_spmv": (C.c_int, [_P, C.c_int32, C.c_int32, _D]),
    "pdfp_device_bytes": (C.c_int64, [_P]),
    "pdfp_set_kernel_path": (C.c_int, [_P, C.c_int32]),
    "pdfp_set_block_graph": (C.c_int, [_P, C.c_int32]),
    "pdfp_set_schedule": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "pdfp_batch_last_error": (C.c_char_p, [_P]),
    "pdfp_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.c_int64, C.c_int64, C.c_int64,
                                    _I64, _I64, _D, C.c_int32, _D, _U8, _U8, _D, _D, _D,
                                    C.c_int32]),
    "pdfp_batch_set_original_primal_residual": (C.c_int, [_P, _D]),
    "pdfp_batch_set_x": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_batch_set_target": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_batch_set_column_scale": (C.c_int, [_P, _D]),
    "pdfp_batch_set_target_original": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_batch_warm_start": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_solve": (C.c_int, [_P, C.POINTER(PdfpParams), _D, _U8, C.POINTER(C.c_int32),
                                   _I64]),
    "pdfp_batch_begin": (C.c_int, [_P, C.c_int32, C.POINTER(PdfpParams), C.c_double]),
    "pdfp_batch_run": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _I64,
                                 C.POINTER(C.c_int32)]),
    "pdfp_batch_get_x": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_batch_get_vector": (C.c_int, [_P, C.c_int32, C.c_int32, _D]),
    "pdfp_batch_get_info": (C.c_int, [_P, C.c_int32, C.POINTER(PdfpInfo)]),
    "pdfp_batch_time_iterations": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_batch_time_kernels": (C.c_int, [_P, C.c_int32, _D, _D]),
    "pdfp_batch_set_schedule": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32]),
    "pdfp_batch_set_block_graph": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_set_cooperative": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_set_column_slices": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_set_slice_min_len": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_set_nontemporal": (C.c_int, [_P, C.c_int32]),
    "pdfp_batch_set_kernel_path": (C.c_int, [_P, C.c_int32]),
    "pdfp_pack_last_error": (C.c_char_p, [_P]),
    "pdfp_pack_create": (C.c_int, [C.POINTER(_P), C.c_int32, _I64, _I64, C.POINTER(_I64),
                                   C.POINTER(_I64), C.POINTER(_D), C.c_int32, C.POINTER(_D),
                                   C.POINTER(_U8), C.POINTER(_U8), C.POINTER(_D), C.POINTER(_D),
                                   C.POINTER(_D), C.c_int32]),
    "pdfp_pack_set_original_primal_residual": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_pack_set_x": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_pack_set_target": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_pack_warm_start": (C.c_int, [_P, C.c_int32]),
    "pdfp_pack_solve": (C.c_int, [_P, C.POINTER(PdfpParams), _D, _U8, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int64)]),
    "pdfp_pack_begin": (C.c_int, [_P, C.c_int32, C.POINTER(PdfpParams), C.c_double]),
    "pdfp_pack_run": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]),
    "pdfp_pack_get_x": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_pack_get_vector": (C.c_int, [_P, C.c_int32, C.c_int32, _D]),
    "pdfp_pack_set_block_graph": (C.c_int, [_P, C.c_int32]),
    "pdfp_pack_set_kernel_path": (C.c_int, [_P, C.c_int32]),
    "pdfp_pack_get_info": (C.c_int, [_P, C.c_int32, C.POINTER(PdfpInfo)]),
    "pdfp_pack_time_iterations": (C.c_int, [_P, C.c_int32, _D]),
    "pdfp_pack_time_kernels": (C.c_int, [_P, C.c_int32, _D, _D]),
    "pdfp_pack_destroy": (None, [_P]),
    "pdfp_batch_destroy": (None, [_P]),
    "pdfp_scale_matrix": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, _I64, _I64, _D, C.c_int32,
                                    C.c_int32, C.c_double, _D, _D, _D, C.c_int32]),
    "pdfp_scale_last_error": (C.c_char_p, []),
    "pdfp_rounder_create": (C.c_int, [C.POINTER(_P), C.c_int64, C.c_int64, _I64, _I64, _D,
                                      C.c_int32, _D, _U8, _D, _D, _U8, C.c_int64]),
    "pdfp_rounder_set_direction": (C.c_int, [_P, _U8]),
    "pdfp_round_sequential": (C.c_int, [_P, _D, _I64, C.c_int64, _D]),
    "pdfp_round_guarded": (C.c_int, [_P, _D, _I64, C.c_int64, C.c_int32, _D,
                                     C.POINTER(C.c_int32)]),
    "pdfp_rounder_last_error": (C.c_char_p, [_P]),
    "pdfp_rounder_destroy": (None, [_P]),
    "pdfp_checker_create": (C.c_int, [C.POINTER(_P), C.c_int64, C.c_int64, _I64, _I64, _D,
                                      C.c_int32, _D, _U8, C.c_int32]),
    "pdfp_checker_max_violation": (C.c_int, [_P, _D, _D]),
    "pdfp_checker_last_error": (C.c_char_p, [_P]),
    "pdfp_checker_destroy": (None, [_P]),
    # device pointers throughout; the last argument is a hipStream_t
    "pdfp_max_violation_noalloc": (C.c_int, [C.c_int64, C.c_int64, _P, _P, _P, _P, _P, _P,
                                             _P, _P]),
}

# libpdfp_probe (include/pdfp_probe.h): measurement only, loaded by bench.py.
PROBE_SIGNATURES = {
    "pdfp_gather_ceiling": (C.c_int, [C.c_int64, C.c_int32, C.c_int64, C.c_int32, _D]),
    "pdfp_probe_last_error": (C.c_char_p, []),
}

_lib = None


def load() -> C.CDLL:
    """Load libpdfp_hip.so from the package directory (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pdfp_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue  # calling it raises AttributeError; the export test names it
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_I64)


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(_U8)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class LibraryError(RuntimeError):
    pass


_probe = None


def load_probe() -> C.CDLL:
    """Load libpdfp_probe.so (the benchmark's gather probe)."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_PATH):
            raise ImportError(f"{PROBE_PATH} is missing: build it with `python -m pdfp_amd.build`")
        lib = C.CDLL(PROBE_PATH)
        for name, (res, args) in PROBE_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _probe = lib
    return _probe


def gather_ceiling_us(table_bytes: int, record_bytes: int, n_gathers: int,
                      device: int = 0) -> float:
    """Microseconds per pass of n_gathers random gathers of record_bytes-byte
    records from a table_bytes-byte device table (pdfp_gather_ceiling)."""
    lib = load_probe()
    us = C.c_double()
    st = lib.pdfp_gather_ceiling(int(table_bytes), int(record_bytes), int(n_gathers), device,
                                 C.byref(us))
    if st != 0:
        raise LibraryError(f"libpdfp_probe error {st}: {lib.pdfp_probe_last_error().decode()}")
    return us.value


def check(status: int, err_fn, ctx) -> None:
    if status != 0:
        msg = err_fn(ctx)
        raise LibraryError(f"libpdfp_hip error {status}: {msg.decode() if msg else '?'}")
