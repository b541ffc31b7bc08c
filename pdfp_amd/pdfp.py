"""The PDFP projector on the GPU: the host mirror of src/pdfp.jl.

``PDFPState`` owns a libpdfp_hip context: the scaled CSR A and Aᵀ, the
target, the weights and every iterate live in HBM for the whole pump run;
each projection replaces only the target and warm-starts from the previous
iterate.  The reference's functions map one to one:

=================================  ==========================================
reference (src/pdfp.jl)            here
=================================  ==========================================
``PDFPState(A, b, ...; use_gpu)``  ``PDFPState(A, b, ..., precision=...)``
``original_primal_residual!``      ``original_primal_residual(s, D_r)``
``copyto!(s.current_x, x)``        ``s.set_x(x)``  (current and initial x)
``set_target!``                    ``set_target(s, x̃)``
``warm_start!``                    ``warm_start(s)``
``pdfp_solve!``                    ``pdfp_solve(s, p, primal_weight)``
``Array(s.pdhg_x)``                ``s.pdhg_x``
``s.total_count``                  ``s.total_count``
=================================  ==========================================

Termination reasons are the reference's Symbols as strings ("OPTIMAL",
"ITERATION_LIMIT", "TIME_LIMIT").
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, asdict

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, dptr, f64, i64ptr, u8ptr


@dataclass
class RestartParameters:
    """src/pdlp.jl:10-18."""

    artificial_restart_threshold: float = 0.36
    sufficient_reduction_for_restart: float = 0.2
    necessary_reduction_for_restart: float = 0.5
    k_p: float = 0.99
    k_i: float = 0.01
    k_d: float = 0.0
    i_smooth: float = 0.3


@dataclass
class SolverParameters:
    """src/pdlp.jl:21-31 (the defaults are cuPDLPx's)."""

    eps_optimal_relative: float = 1e-4
    eps_feasible_relative: float = 1e-4
    iteration_limit: int = 100_000
    time_limit: float = 3600.0
    termination_evaluation_frequency: int = 200
    reflection_coefficient: float = 1.0
    sv_max_iterations: int = 5000
    sv_tolerance: float = 1e-4
    restart: RestartParameters = field(default_factory=RestartParameters)

    def to_c(self) -> _lib.PdfpParams:
        p = _lib.PdfpParams()
        p.eps_optimal_relative = self.eps_optimal_relative
        p.eps_feasible_relative = self.eps_feasible_relative
        p.iteration_limit = int(self.iteration_limit)
        p.time_limit = float(self.time_limit)
        p.termination_evaluation_frequency = int(self.termination_evaluation_frequency)
        p.sv_max_iterations = int(self.sv_max_iterations)
        p.reflection_coefficient = self.reflection_coefficient
        p.sv_tolerance = self.sv_tolerance
        for k, v in asdict(self.restart).items():
            setattr(p, k, float(v))
        return p


def _csc_arrays(A: sp.csc_matrix):
    A = sp.csc_matrix(A, dtype=np.float64)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, dtype=np.int64),
            np.ascontiguousarray(A.indices, dtype=np.int64),
            np.ascontiguousarray(A.data, dtype=np.float64))


class PDFPState:
    """Projector state on the device (src/pdfp.jl:15-76).

    ``precision``: "fp64" (the reference's FP64 arithmetic and launch
    structure) or "fp32" (FP32 data and iterates, fused launches, FP64
    reductions and polish).
    """

    def __init__(self, A: sp.csc_matrix, b, is_eq, is_int, var_lb, var_ub, w, *,
                 precision: str = "fp32", device: int = 0):
        lib = _lib.load()
        self._lib = lib
        m, n = A.shape
        self.m, self.n = int(m), int(n)
        colptr, rowval, nzval = _csc_arrays(A)
        self._b = f64(b)
        eq = np.ascontiguousarray(is_eq, dtype=np.uint8)
        isint = np.ascontiguousarray(is_int, dtype=np.uint8)
        lb, ub, ww = f64(var_lb), f64(var_ub), f64(w)
        for name, v, ln in (("b", self._b, m), ("is_eq", eq, m), ("is_int", isint, n),
                            ("var_lb", lb, n), ("var_ub", ub, n), ("w", ww, n)):
            if len(v) != ln:
                raise ValueError(f"{name} has length {len(v)}, expected {ln}")
        if precision not in _lib.PRECISION:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        ctx = C.c_void_p()
        st = lib.pdfp_create(C.byref(ctx), m, n, len(nzval), i64ptr(colptr), i64ptr(rowval),
                             dptr(nzval), 0, dptr(self._b), u8ptr(eq), u8ptr(isint), dptr(lb),
                             dptr(ub), dptr(ww), _lib.PRECISION[precision], device)
        check(st, lib.pdfp_last_error, None)
        self._ctx = ctx
        self.termination_reason = "UNSPECIFIED"

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdfp_destroy(ctx)
            self._ctx = None

    def close(self):
        self.__del__()

    def _check(self, st):
        check(st, self._lib.pdfp_last_error, self._ctx)

    # -- the reference's mutators --------------------------------------------
    def set_x(self, x_scaled) -> None:
        """copyto!(current_x, x); copyto!(initial_x, x)  (src/pump.jl:305-306)."""
        x = f64(x_scaled)
        self._check(self._lib.pdfp_set_x(self._ctx, dptr(x)))

    def set_dual(self, y, z) -> None:
        """Current and initial y and z (no counterpart in the reference: the
        FP64 polish starts from an FP32 context's duals)."""
        yy, zz = f64(y), f64(z)
        if len(yy) != self.n or len(zz) != self.m:
            raise ValueError("y must have n entries and z m")
        self._check(self._lib.pdfp_set_dual(self._ctx, dptr(yy), dptr(zz)))

    def info(self) -> _lib.PdfpInfo:
        out = _lib.PdfpInfo()
        self._check(self._lib.pdfp_get_info(self._ctx, C.byref(out)))
        return out

    @property
    def pdhg_x(self) -> np.ndarray:
        """Array(s.pdhg_x): the projection in scaled units (FP64; polished in
        the fp32 form)."""
        out = np.empty(self.n)
        self._check(self._lib.pdfp_get_x(self._ctx, dptr(out)))
        return out

    def vector(self, name: str) -> np.ndarray:
        ids = {"pdhg_x": 0, "pdhg_y": 1, "pdhg_z": 2, "current_x": 3, "current_y": 4,
               "current_z": 5, "dual_slack": 6, "x_tilde": 7}
        which = ids[name]
        out = np.empty(self.m if name in ("pdhg_z", "current_z") else self.n)
        self._check(self._lib.pdfp_get_vector(self._ctx, which, dptr(out)))
        return out

    @property
    def total_count(self) -> int:
        return int(self.info().total_count)

    def max_singular_value(self, max_iterations: int = 5000, tolerance: float = 1e-4) -> float:
        out = C.c_double()
        self._check(self._lib.pdfp_max_singular_value(self._ctx, max_iterations, tolerance,
                                                       C.byref(out)))
        return out.value

    def spmv(self, x, transpose: bool = False) -> np.ndarray:
        x = f64(x)
        out = np.empty(self.n if transpose else self.m)
        self._check(self._lib.pdfp_spmv(self._ctx, int(transpose), dptr(x), dptr(out)))
        return out

    def time_iterations(self, iterations: int) -> float:
        ms = C.c_double()
        self._check(self._lib.pdfp_time_iterations(self._ctx, iterations, C.byref(ms)))
        return ms.value

    def time_kernels(self, iterations: int) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.pdfp_time_kernels(self._ctx, iterations, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_spmv(self, iterations: int, transpose: bool = False) -> float:
        """Mean ms of a plain SpMV y = A x (or Aᵀ x) over the resident operator."""
        ms = C.c_double()
        self._check(self._lib.pdfp_time_spmv(self._ctx, int(transpose), iterations,
                                             C.byref(ms)))
        return ms.value

    @property
    def device_bytes(self) -> int:
        return int(self._lib.pdfp_device_bytes(self._ctx))

    def set_kernel_path(self, path: str) -> None:
        """"auto" (the default): the FP32 form runs each block of iterations
        in one workgroup when the operators fit in LDS, the FP64 form the
        reference's five launches per iteration; "general" always runs the
        general passes; "one-workgroup" runs blocks (and the power iteration)
        in one workgroup in either form when the instance fits.  Set it
        before the first solve."""
        codes = {"auto": 0, "general": 1, "one-workgroup": 2}
        if path not in codes:
            raise ValueError("kernel path must be 'auto', 'general' or 'one-workgroup'")
        self._check(self._lib.pdfp_set_kernel_path(self._ctx, codes[path]))

    def set_block_graph(self, on: bool) -> None:
        """FP32 form, general passes: replay each block of iterations (with
        its evaluation) as one captured HIP graph (the default), or launch
        every kernel on the stream.  Both give identical results."""
        self._check(self._lib.pdfp_set_block_graph(self._ctx, int(bool(on))))

    def set_schedule(self, npb_a: int, npb_at: int, persist_a: int = 0,
                     persist_at: int = 0) -> None:
        """Replace the autotuned schedule of the A and Aᵀ passes."""
        self._check(self._lib.pdfp_set_schedule(self._ctx, npb_a, npb_at, persist_a,
                                                persist_at))


def original_primal_residual(s: PDFPState, D_r) -> PDFPState:
    """src/pdfp.jl:85-90."""
    d = f64(D_r)
    s._check(s._lib.pdfp_set_original_primal_residual(s._ctx, dptr(d)))
    return s


def set_target(s: PDFPState, x_tilde_scaled) -> PDFPState:
    """src/pdfp.jl:98-101."""
    x = f64(x_tilde_scaled)
    s._check(s._lib.pdfp_set_target(s._ctx, dptr(x)))
    s.termination_reason = "UNSPECIFIED"
    return s


def set_column_scale(s: PDFPState, D_c) -> PDFPState:
    """The column scaling D_c (n positive values) for set_target_original."""
    d = f64(D_c)
    if d.shape != (s.n,):
        raise ValueError(f"D_c has shape {d.shape}, expected ({s.n},)")
    s._check(s._lib.pdfp_set_column_scale(s._ctx, dptr(d)))
    return s


def set_target_original(s: PDFPState, x_tilde) -> PDFPState:
    """The pump's ``set_target!(projector, x̃ ./ D_c)`` (src/pump.jl:331) with
    x̃ in original units and the division on the device (after
    set_column_scale); equal, bit for bit, to set_target(s, x̃ / D_c)."""
    x = f64(x_tilde)
    if x.shape != (s.n,):
        raise ValueError(f"x_tilde has shape {x.shape}, expected ({s.n},)")
    s._check(s._lib.pdfp_set_target_original(s._ctx, dptr(x)))
    s.termination_reason = "UNSPECIFIED"
    return s


def warm_start(s: PDFPState) -> PDFPState:
    """src/pdfp.jl:104-114."""
    s._check(s._lib.pdfp_warm_start(s._ctx))
    s.termination_reason = "UNSPECIFIED"
    return s


def pdfp_solve(s: PDFPState, p: SolverParameters, primal_weight: float) -> str:
    """src/pdfp.jl:252-288.  Returns the termination reason."""
    term = C.c_int32()
    total = C.c_int64()
    cp = p.to_c()
    s._check(s._lib.pdfp_solve(s._ctx, C.byref(cp), float(primal_weight), C.byref(term),
                               C.byref(total)))
    s.termination_reason = _lib.TERMINATION[term.value]
    return s.termination_reason


@dataclass
class ProjectorProblem:
    """The projector's scaled data, as the pump builds it
    (src/pump.jl:295-303; test/runtests.jl:157-166)."""

    A: sp.csc_matrix
    b: np.ndarray
    is_eq: np.ndarray
    is_int: np.ndarray
    var_lb: np.ndarray
    var_ub: np.ndarray
    w: np.ndarray
    D_r: np.ndarray
    D_c: np.ndarray


def projector_problem(milp, *, scaling: str = "device", device: int = 0) -> ProjectorProblem:
    """Scale ``milp`` for the projector: no bound/objective rescaling, the
    ℓ1 weights are D_c on I and 1 elsewhere (src/pump.jl:295-301).
    ``scaling``: where the matrix is equilibrated, "device" or "host"."""
    from .preconditioner import scale_problem

    n = len(milp.c)
    m = len(milp.b)
    is_eq = milp.row_sense == "eq"
    is_int = np.zeros(n, dtype=bool)
    is_int[milp.int_idx] = True
    con_lb = np.where(is_eq, milp.b, -np.inf) if m else np.zeros(0)
    spr = scale_problem(milp.A, np.zeros(n), milp.l.copy(), milp.u.copy(), con_lb, milp.b.copy(),
                        bound_objective_rescaling=False, backend=scaling, device=device)
    w = np.ones(n)
    w[milp.int_idx] = spr.D_c[milp.int_idx]
    return ProjectorProblem(A=spr.A, b=spr.con_ub, is_eq=is_eq, is_int=is_int, var_lb=spr.var_lb,
                            var_ub=spr.var_ub, w=w, D_r=spr.D_r, D_c=spr.D_c)


def project(milp, x_tilde, *, tol: float = 1e-8, precision: str = "fp64",
            iteration_limit: int = 1_000_000):
    """The test helper of the reference (test/runtests.jl:12-27): the ℓ1
    projection of ``x_tilde`` onto the LP relaxation of ``milp``, from zero,
    with primal weight 1.  Returns (x in original units, termination)."""
    pp = projector_problem(milp)
    s = PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                  precision=precision)
    set_target(s, np.asarray(x_tilde, dtype=np.float64) / pp.D_c)
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=iteration_limit)
    status = pdfp_solve(s, p, 1.0)
    x = pp.D_c * s.pdhg_x
    s.close()
    return x, status
