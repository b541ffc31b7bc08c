"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/liboracle.so, the plain-C FP64 restatement of
the reference's CPU path (src/preconditioner.jl, src/pdfp.jl, src/pdlp.jl).
Only the tests, the smoke check and the CPU-baseline leg of the benchmark
use it, as the checker.  ``OraclePDFPState`` has the same methods as
``pdfp_amd.pdfp.PDFPState`` so the same test code drives both.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)


class OrcParams(C.Structure):
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


class OrcInfo(C.Structure):
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
    ]


class OrcPdlpInfo(C.Structure):
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
    ]


TERMINATION = {0: "UNSPECIFIED", 1: "OPTIMAL", 2: "ITERATION_LIMIT", 3: "TIME_LIMIT",
               4: "PRIMAL_INFEASIBLE", 5: "DUAL_INFEASIBLE"}

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "orc_default_params": (None, [C.POINTER(OrcParams)]),
        "orc_scale_problem": (None, [C.c_int64, C.c_int64, _I64, _I64, _D, _D, _D, _D, _D, _D,
                                     C.c_int32, C.c_double, C.c_int32, _D, _D, _D, _D]),
        "orc_start_vector": (None, [C.c_int64, C.c_uint64, _D]),
        "orc_max_singular_value": (C.c_double, [C.c_int64, C.c_int64, _I64, _I64, _D, C.c_int32,
                                                C.c_double]),
        "orc_spmv": (None, [C.c_int64, C.c_int64, _I64, _I64, _D, C.c_int32, _D, _D]),
        "orc_set_threads": (None, [C.c_int32]),
        "orc_pdfp_create": (P, [C.c_int64, C.c_int64, _I64, _I64, _D, _D, _U8, _U8, _D, _D, _D]),
        "orc_pdfp_destroy": (None, [P]),
        "orc_pdfp_original_primal_residual": (None, [P, _D]),
        "orc_pdfp_set_x": (None, [P, _D]),
        "orc_pdfp_set_dual": (None, [P, _D, _D]),
        "orc_pdfp_warm_start": (None, [P]),
        "orc_pdfp_set_target": (None, [P, _D]),
        "orc_pdfp_solve": (C.c_int32, [P, C.POINTER(OrcParams), C.c_double]),
        "orc_pdfp_get_vector": (None, [P, C.c_int32, _D]),
        "orc_pdfp_get_info": (None, [P, C.POINTER(OrcInfo)]),
        "orc_pdfp_residual_of": (None, [P, _D, _D, _D, _D, _D]),
        "orc_pdfp_initialize": (None, [P, C.POINTER(OrcParams), C.c_double]),
        "orc_pdlp_create": (P, [C.c_int64, C.c_int64, _I64, _I64, _D, _D, _D, _D, _D, _D,
                                C.c_double, C.c_double]),
        "orc_pdlp_destroy": (None, [P]),
        "orc_pdlp_cupdlpx_termination": (None, [P, _D, _D]),
        "orc_pdlp_solve": (C.c_int32, [P, C.POINTER(OrcParams)]),
        "orc_pdlp_get_vector": (None, [P, C.c_int32, _D]),
        "orc_pdlp_get_info": (None, [P, C.POINTER(OrcPdlpInfo)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _csc(A):
    import scipy.sparse as sp

    A = sp.csc_matrix(A, dtype=np.float64)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, dtype=np.int64),
            np.ascontiguousarray(A.indices, dtype=np.int64),
            np.ascontiguousarray(A.data, dtype=np.float64))


def params(p=None) -> OrcParams:
    """OrcParams from a pdfp_amd SolverParameters (or the defaults)."""
    out = OrcParams()
    load().orc_default_params(C.byref(out))
    if p is not None:
        out.eps_optimal_relative = p.eps_optimal_relative
        out.eps_feasible_relative = p.eps_feasible_relative
        out.iteration_limit = int(p.iteration_limit)
        out.time_limit = float(p.time_limit)
        out.termination_evaluation_frequency = int(p.termination_evaluation_frequency)
        out.sv_max_iterations = int(p.sv_max_iterations)
        out.reflection_coefficient = p.reflection_coefficient
        out.sv_tolerance = p.sv_tolerance
        r = p.restart
        out.artificial_restart_threshold = r.artificial_restart_threshold
        out.sufficient_reduction_for_restart = r.sufficient_reduction_for_restart
        out.necessary_reduction_for_restart = r.necessary_reduction_for_restart
        out.k_p, out.k_i, out.k_d, out.i_smooth = r.k_p, r.k_i, r.k_d, r.i_smooth
    return out


def scale_problem(A, c, var_lb, var_ub, con_lb, con_ub, *, ruiz_iterations=10,
                  pock_chambolle_alpha=1.0, bound_objective_rescaling=True):
    """src/preconditioner.jl:26-99; returns a dict of the ScaledProblem fields."""
    lib = load()
    colptr, rowval, vals = _csc(A)
    vals = vals.copy()
    m, n = A.shape
    c, lb, ub, clb, cub = (_f64(v).copy() for v in (c, var_lb, var_ub, con_lb, con_ub))
    D_r, D_c = np.empty(m), np.empty(n)
    bs, os_ = C.c_double(), C.c_double()
    lib.orc_scale_problem(m, n, colptr.ctypes.data_as(_I64), rowval.ctypes.data_as(_I64),
                          _d(vals), _d(c), _d(lb), _d(ub), _d(clb), _d(cub), ruiz_iterations,
                          pock_chambolle_alpha, int(bound_objective_rescaling), _d(D_r), _d(D_c),
                          C.byref(bs), C.byref(os_))
    import scipy.sparse as sp

    As = sp.csc_matrix((vals, rowval, colptr), shape=(m, n))
    return dict(A=As, c=c, var_lb=lb, var_ub=ub, con_lb=clb, con_ub=cub, D_r=D_r, D_c=D_c,
                bound_scale=bs.value, objective_scale=os_.value)


def start_vector(m: int, seed: int = 20240517) -> np.ndarray:
    out = np.empty(m)
    load().orc_start_vector(m, seed, _d(out))
    return out


def max_singular_value(A, max_iterations=5000, tolerance=1e-4) -> float:
    colptr, rowval, vals = _csc(A)
    m, n = A.shape
    return load().orc_max_singular_value(m, n, colptr.ctypes.data_as(_I64),
                                         rowval.ctypes.data_as(_I64), _d(vals), max_iterations,
                                         tolerance)


def set_threads(n: int) -> None:
    """Host threads for the projector's step: its two products and three
    elementwise updates (1, the reference's single-threaded CPU path, by
    default); the iterates stay bit-identical."""
    load().orc_set_threads(int(n))


def spmv(A, x, transpose=False) -> np.ndarray:
    colptr, rowval, vals = _csc(A)
    m, n = A.shape
    x = _f64(x)
    out = np.empty(n if transpose else m)
    load().orc_spmv(m, n, colptr.ctypes.data_as(_I64), rowval.ctypes.data_as(_I64), _d(vals),
                    int(transpose), _d(x), _d(out))
    return out


class OraclePDFPState:
    """FP64 CPU PDFPState with the interface of pdfp_amd.pdfp.PDFPState."""

    def __init__(self, A, b, is_eq, is_int, var_lb, var_ub, w, **_ignored):
        lib = load()
        self._lib = lib
        colptr, rowval, vals = _csc(A)
        self.m, self.n = A.shape
        eq = np.ascontiguousarray(is_eq, dtype=np.uint8)
        ii = np.ascontiguousarray(is_int, dtype=np.uint8)
        self._ctx = lib.orc_pdfp_create(self.m, self.n, colptr.ctypes.data_as(_I64),
                                        rowval.ctypes.data_as(_I64), _d(vals), _d(_f64(b)),
                                        eq.ctypes.data_as(_U8), ii.ctypes.data_as(_U8),
                                        _d(_f64(var_lb)), _d(_f64(var_ub)), _d(_f64(w)))
        self.termination_reason = "UNSPECIFIED"
        self.precision = "oracle"

    def __del__(self):
        if getattr(self, "_ctx", None):
            self._lib.orc_pdfp_destroy(self._ctx)
            self._ctx = None

    def close(self):
        self.__del__()

    def set_x(self, x):
        self._lib.orc_pdfp_set_x(self._ctx, _d(_f64(x)))

    def set_dual(self, y, z):
        self._lib.orc_pdfp_set_dual(self._ctx, _d(_f64(y)), _d(_f64(z)))

    def info(self) -> OrcInfo:
        out = OrcInfo()
        self._lib.orc_pdfp_get_info(self._ctx, C.byref(out))
        return out

    def vector(self, name: str) -> np.ndarray:
        ids = {"pdhg_x": 0, "pdhg_y": 1, "pdhg_z": 2, "current_x": 3, "current_y": 4,
               "current_z": 5, "dual_slack": 6, "x_tilde": 7}
        out = np.empty(self.m if name in ("pdhg_z", "current_z") else self.n)
        self._lib.orc_pdfp_get_vector(self._ctx, ids[name], _d(out))
        return out

    @property
    def pdhg_x(self) -> np.ndarray:
        return self.vector("pdhg_x")

    @property
    def total_count(self) -> int:
        return int(self.info().total_count)

    def residual_of(self, x, y, z, dual_slack) -> np.ndarray:
        """compute_residual! (src/pdfp.jl:200-218) on a given point: relative
        primal, relative dual, relative ℓ1 gap, ℓ1 distance."""
        out = np.empty(4)
        self._lib.orc_pdfp_residual_of(self._ctx, _d(_f64(x)), _d(_f64(y)), _d(_f64(z)),
                                       _d(_f64(dual_slack)), _d(out))
        return out

    def initialize(self, p=None, primal_weight=1.0):
        cp = params(p)
        self._lib.orc_pdfp_initialize(self._ctx, C.byref(cp), float(primal_weight))


def original_primal_residual(s: OraclePDFPState, D_r):
    s._lib.orc_pdfp_original_primal_residual(s._ctx, _d(_f64(D_r)))
    return s


def set_target(s: OraclePDFPState, x):
    s._lib.orc_pdfp_set_target(s._ctx, _d(_f64(x)))
    s.termination_reason = "UNSPECIFIED"
    return s


def warm_start(s: OraclePDFPState):
    s._lib.orc_pdfp_warm_start(s._ctx)
    return s


def pdfp_solve(s: OraclePDFPState, p, primal_weight: float) -> str:
    cp = params(p)
    t = s._lib.orc_pdfp_solve(s._ctx, C.byref(cp), float(primal_weight))
    s.termination_reason = TERMINATION[t]
    return s.termination_reason


class OraclePDLPState:
    """FP64 CPU PDLPState (src/pdlp.jl:33-96) on a scaled problem dict."""

    def __init__(self, sp_: dict):
        lib = load()
        self._lib = lib
        A = sp_["A"]
        colptr, rowval, vals = _csc(A)
        self.m, self.n = A.shape
        self._ctx = lib.orc_pdlp_create(self.m, self.n, colptr.ctypes.data_as(_I64),
                                        rowval.ctypes.data_as(_I64), _d(vals),
                                        _d(_f64(sp_["c"])), _d(_f64(sp_["var_lb"])),
                                        _d(_f64(sp_["var_ub"])), _d(_f64(sp_["con_lb"])),
                                        _d(_f64(sp_["con_ub"])), float(sp_["bound_scale"]),
                                        float(sp_["objective_scale"]))

    def __del__(self):
        if getattr(self, "_ctx", None):
            self._lib.orc_pdlp_destroy(self._ctx)
            self._ctx = None

    def cupdlpx_termination(self, D_r, D_c):
        self._lib.orc_pdlp_cupdlpx_termination(self._ctx, _d(_f64(D_r)), _d(_f64(D_c)))

    def solve(self, p) -> str:
        cp = params(p)
        return TERMINATION[self._lib.orc_pdlp_solve(self._ctx, C.byref(cp))]

    def vector(self, name: str) -> np.ndarray:
        ids = {"pdhg_primal": 0, "dual_slack": 1, "pdhg_dual": 2, "current_primal": 3,
               "current_dual": 4}
        out = np.empty(self.m if name in ("pdhg_dual", "current_dual") else self.n)
        self._lib.orc_pdlp_get_vector(self._ctx, ids[name], _d(out))
        return out

    def info(self) -> OrcPdlpInfo:
        out = OrcPdlpInfo()
        self._lib.orc_pdlp_get_info(self._ctx, C.byref(out))
        return out


def solve_lp_relaxation(milp, *, tolerance=1e-2, time_limit=3600.0, cupdlpx_termination=False):
    """src/pdlp.jl:381-393 on the oracle: (x, dual_slack, status, iterations)."""
    from pdfp_amd.pdfp import SolverParameters

    m = len(milp.b)
    con_lb = np.where(milp.row_sense == "le", -np.inf, milp.b)
    sp_ = scale_problem(milp.A, milp.c.copy(), milp.l.copy(), milp.u.copy(), con_lb,
                        milp.b.copy())
    s = OraclePDLPState(sp_)
    if cupdlpx_termination:
        s.cupdlpx_termination(sp_["D_r"], sp_["D_c"])
    p = SolverParameters(iteration_limit=5_000_000, eps_optimal_relative=tolerance,
                         eps_feasible_relative=tolerance, time_limit=time_limit)
    status = s.solve(p)
    x = (sp_["D_c"] * s.vector("pdhg_primal")) / sp_["bound_scale"]
    return x, s.vector("dual_slack"), status, int(s.info().total_count)
