"""Diagonal preconditioning of the PDHG solvers (host, FP64).

``scale_problem`` restates src/preconditioner.jl:26-99: 10 Ruiz rounds, one
Pock–Chambolle pass with α = 1 and, optionally, the bound/objective
rescaling.  The scalings D_r (rows) and D_c (columns) change variables to
x = D_c ∘ x̂: the objective becomes (D_c ∘ c)ᵀ x̂, the rows read
D_r ∘ con_lb ≤ diag(D_r) A diag(D_c) x̂ ≤ D_r ∘ con_ub and the column bounds
become var_lb / D_c ≤ x̂ ≤ var_ub / D_c.  The arithmetic is max, sqrt, divide and multiply, in the
reference's order, so the result is bit-identical to the reference's CPU
code; the Pock–Chambolle sums accumulate in the CSC's column-major order.

The matrix part (the Ruiz rounds and the Pock–Chambolle pass, ≈22 passes over
the nonzeros) runs on the device by default (``pdfp_scale_matrix``, one
launch per pass); ``backend="host"`` runs the same arithmetic in numpy
(np.bincount adds sequentially in input order).  Both are bit-identical to
the oracle.  The vector part (src/preconditioner.jl:65-99) is O(m + n) and
runs on the host.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

SCALING_EPSILON = 1e-12


@dataclass
class ScaledProblem:
    """src/preconditioner.jl:13-24."""

    A: sp.csc_matrix
    c: np.ndarray
    var_lb: np.ndarray
    var_ub: np.ndarray
    con_lb: np.ndarray
    con_ub: np.ndarray
    D_r: np.ndarray
    D_c: np.ndarray
    bound_scale: float
    objective_scale: float


def _inv_sqrt_or_one(v: np.ndarray) -> np.ndarray:
    out = np.ones_like(v)
    big = ~(v < SCALING_EPSILON)
    out[big] = 1.0 / np.sqrt(v[big])
    return out


def _col_of_entry(A: sp.csc_matrix) -> np.ndarray:
    return np.repeat(np.arange(A.shape[1], dtype=np.int64), np.diff(A.indptr))


def _segment_max(vals: np.ndarray, indptr: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros(n)
    nz = np.diff(indptr) > 0
    if vals.size:
        starts = indptr[:-1][nz]
        out[nz] = np.maximum.reduceat(vals, starts)
    return np.maximum(out, 0.0)


def _scale_matrix_host(A: sp.csc_matrix, ruiz_iterations: int, alpha: float):
    """src/preconditioner.jl:33-63 in numpy: returns (scaled values in CSC
    order, D_r, D_c)."""
    m, n = A.shape
    rows = A.indices.astype(np.int64)
    cols = _col_of_entry(A)
    vals = A.data.copy()
    D_r = np.ones(m)
    D_c = np.ones(n)
    # entries in row-major order, for the row maxima (max is order-free)
    perm = np.argsort(rows, kind="stable")
    row_ptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=m), out=row_ptr[1:])
    for _ in range(ruiz_iterations):
        av = np.abs(vals)
        row_max = _segment_max(av[perm], row_ptr, m)
        col_max = _segment_max(av, A.indptr, n)
        rscale = _inv_sqrt_or_one(row_max)
        cscale = _inv_sqrt_or_one(col_max)
        vals *= rscale[rows] * cscale[cols]
        D_r *= rscale
        D_c *= cscale

    av = np.abs(vals)
    # v^α with α == 1 is v itself; np.power(v, 1.0) returns v exactly.
    row_sum = np.bincount(rows, weights=np.power(av, alpha), minlength=m)
    col_sum = np.bincount(cols, weights=np.power(av, 2.0 - alpha), minlength=n)
    rscale = _inv_sqrt_or_one(row_sum.astype(np.float64))
    cscale = _inv_sqrt_or_one(col_sum.astype(np.float64))
    vals *= rscale[rows] * cscale[cols]
    D_r *= rscale
    D_c *= cscale
    return vals, D_r, D_c


def _scale_matrix_device(A: sp.csc_matrix, ruiz_iterations: int, alpha: float, device: int):
    """The matrix part on the device (pdfp_scale_matrix)."""
    from . import _lib

    lib = _lib.load()
    m, n = A.shape
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
    nzval = np.ascontiguousarray(A.data, dtype=np.float64)
    vals = np.empty(len(nzval))
    D_r = np.empty(m)
    D_c = np.empty(n)
    st = lib.pdfp_scale_matrix(m, n, len(nzval), _lib.i64ptr(colptr), _lib.i64ptr(rowval),
                               _lib.dptr(nzval), 0, int(ruiz_iterations), float(alpha),
                               _lib.dptr(vals), _lib.dptr(D_r), _lib.dptr(D_c), device)
    if st != 0:
        raise _lib.LibraryError(f"libpdfp_hip error {st}: "
                                f"{lib.pdfp_scale_last_error().decode()}")
    return vals, D_r, D_c


def scale_problem(A: sp.csc_matrix, c, var_lb, var_ub, con_lb, con_ub, *,
                  ruiz_iterations: int = 10, pock_chambolle_alpha: float = 1.0,
                  bound_objective_rescaling: bool = True, backend: str = "device",
                  device: int = 0) -> ScaledProblem:
    """src/preconditioner.jl:26-99.  ``backend``: "device" (the matrix part
    on the GPU) or "host"."""
    A = sp.csc_matrix(A, dtype=np.float64, copy=True)
    A.sort_indices()
    m, n = A.shape
    c = np.asarray(c, dtype=np.float64).copy()
    var_lb = np.asarray(var_lb, dtype=np.float64)
    var_ub = np.asarray(var_ub, dtype=np.float64)
    con_lb = np.asarray(con_lb, dtype=np.float64)
    con_ub = np.asarray(con_ub, dtype=np.float64)
    if backend == "device":
        vals, D_r, D_c = _scale_matrix_device(A, ruiz_iterations, pock_chambolle_alpha, device)
    elif backend == "host":
        vals, D_r, D_c = _scale_matrix_host(A, ruiz_iterations, pock_chambolle_alpha)
    else:
        raise ValueError("backend must be 'device' or 'host'")

    c_s = c * D_c
    with np.errstate(invalid="ignore"):
        var_lb_s = np.where(np.isfinite(var_lb), var_lb / D_c, var_lb)
        var_ub_s = np.where(np.isfinite(var_ub), var_ub / D_c, var_ub)
        con_lb_s = np.where(np.isfinite(con_lb), con_lb * D_r, con_lb)
        con_ub_s = np.where(np.isfinite(con_ub), con_ub * D_r, con_ub)

    bound_scale = 1.0
    objective_scale = 1.0
    if bound_objective_rescaling:
        finite_lo = np.isfinite(con_lb)
        finite_hi = np.isfinite(con_ub)
        with np.errstate(invalid="ignore"):
            use_lo = finite_lo & (~finite_hi | (np.abs(con_lb_s - con_ub_s) > SCALING_EPSILON))
        # The reference adds lo² then hi² row by row; np.cumsum accumulates
        # sequentially, so interleaving the terms keeps that order (an
        # omitted term is +0.0, which leaves the running sum unchanged).
        terms = np.zeros(2 * m)
        terms[0::2] = np.where(use_lo, con_lb_s * con_lb_s, 0.0)
        terms[1::2] = np.where(finite_hi, con_ub_s * con_ub_s, 0.0)
        bound_sq = float(np.cumsum(terms)[-1]) if m else 0.0
        bound_scale = 1.0 / (np.sqrt(bound_sq) + 1.0)
        c_sq = float(np.cumsum(c_s * c_s)[-1]) if n else 0.0
        objective_scale = 1.0 / (np.sqrt(c_sq) + 1.0)
        con_lb_s = np.where(finite_lo, con_lb_s * bound_scale, con_lb_s)
        con_ub_s = np.where(finite_hi, con_ub_s * bound_scale, con_ub_s)
        var_lb_s = np.where(np.isfinite(var_lb), var_lb_s * bound_scale, var_lb_s)
        var_ub_s = np.where(np.isfinite(var_ub), var_ub_s * bound_scale, var_ub_s)
        c_s = c_s * objective_scale

    A.data = vals
    return ScaledProblem(A=A, c=c_s, var_lb=var_lb_s, var_ub=var_ub_s, con_lb=con_lb_s,
                         con_ub=con_ub_s, D_r=D_r, D_c=D_c, bound_scale=float(bound_scale),
                         objective_scale=float(objective_scale))
