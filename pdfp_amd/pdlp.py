"""The root-LP solver on the GPU: the host mirror of src/pdlp.jl.

=================================  ==========================================
reference (src/pdlp.jl)            here
=================================  ==========================================
``PDLPState(sp; use_gpu)``         ``PDLPState(sp)``
``cupdlpx_termination!``           ``cupdlpx_termination(s, D_r, D_c)``
``pdlp_solve!``                    ``pdlp_solve(s, p)``
``solve_lp_relaxation``            ``solve_lp_relaxation(milp; ...)``
=================================  ==========================================
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr, f64, i64ptr
from .milp import MILP
from .pdfp import SolverParameters, _csc_arrays
from .preconditioner import ScaledProblem, scale_problem


class PDLPState:
    """Root-LP state on the device (src/pdlp.jl:33-96), FP64."""

    def __init__(self, sp: ScaledProblem, device: int = 0):
        lib = _lib.load()
        self._lib = lib
        m, n = sp.A.shape
        self.m, self.n = int(m), int(n)
        colptr, rowval, nzval = _csc_arrays(sp.A)
        ctx = C.c_void_p()
        st = lib.pdlp_create(C.byref(ctx), m, n, len(nzval), i64ptr(colptr), i64ptr(rowval),
                             dptr(nzval), 0, dptr(f64(sp.c)), dptr(f64(sp.var_lb)),
                             dptr(f64(sp.var_ub)), dptr(f64(sp.con_lb)), dptr(f64(sp.con_ub)),
                             float(sp.bound_scale), float(sp.objective_scale), device)
        check(st, lib.pdlp_last_error, None)
        self._ctx = ctx
        self.termination_reason = "UNSPECIFIED"

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdlp_destroy(ctx)
            self._ctx = None

    def close(self):
        self.__del__()

    def _check(self, st):
        check(st, self._lib.pdlp_last_error, self._ctx)

    @property
    def pdhg_primal(self) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self._lib.pdlp_get_primal(self._ctx, dptr(out)))
        return out

    @property
    def dual_slack(self) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self._lib.pdlp_get_dual_slack(self._ctx, dptr(out)))
        return out

    def info(self) -> _lib.PdlpInfo:
        out = _lib.PdlpInfo()
        self._check(self._lib.pdlp_get_info(self._ctx, C.byref(out)))
        return out

    @property
    def total_count(self) -> int:
        return int(self.info().total_count)

    def set_kernel_path(self, path: str) -> None:
        """"auto" runs each block of iterations in one workgroup when the
        instance fits in LDS (the default); "general" always runs the
        general passes."""
        codes = {"auto": 0, "general": 1}
        if path not in codes:
            raise ValueError("kernel path must be 'auto' or 'general'")
        self._check(self._lib.pdlp_set_kernel_path(self._ctx, codes[path]))


def cupdlpx_termination(s: PDLPState, D_r, D_c) -> PDLPState:
    """src/pdlp.jl:105-110."""
    r, c = f64(D_r), f64(D_c)
    s._check(s._lib.pdlp_set_cupdlpx_termination(s._ctx, dptr(r), dptr(c)))
    return s


_cupdlpx_termination = cupdlpx_termination  # the keyword below shadows the name


def pdlp_solve(s: PDLPState, p: SolverParameters) -> str:
    """src/pdlp.jl:325-372."""
    term = C.c_int32()
    total = C.c_int64()
    cp = p.to_c()
    s._check(s._lib.pdlp_solve(s._ctx, C.byref(cp), C.byref(term), C.byref(total)))
    s.termination_reason = _lib.TERMINATION[term.value]
    return s.termination_reason


def solve_lp_relaxation(milp: MILP, *, tolerance: float, time_limit: float,
                        cupdlpx_termination: bool, device: int = 0, scaling: str = "device",
                        kernel_path: str = "auto"):
    """src/pdlp.jl:381-393: the LP relaxation by PDLP.  Returns
    (x in original units, dual_slack in scaled units, status, iterations).
    ``kernel_path``: see PDLPState.set_kernel_path."""
    con_lb = np.where(milp.row_sense == "le", -np.inf, milp.b)
    sp = scale_problem(milp.A, milp.c.copy(), milp.l.copy(), milp.u.copy(), con_lb,
                       milp.b.copy(), backend=scaling, device=device)
    s = PDLPState(sp, device=device)
    s.set_kernel_path(kernel_path)
    if cupdlpx_termination:
        _cupdlpx_termination(s, sp.D_r, sp.D_c)
    p = SolverParameters(iteration_limit=5_000_000, eps_optimal_relative=tolerance,
                         eps_feasible_relative=tolerance, time_limit=time_limit)
    status = pdlp_solve(s, p)
    x = (sp.D_c * s.pdhg_primal) / sp.bound_scale
    out = (x, s.dual_slack, status, s.total_count)
    s.close()
    return out
