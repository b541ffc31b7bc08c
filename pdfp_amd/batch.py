"""Batch of up to 16 projector trajectories of one instance on the device.

The reference runs concurrent trajectories as separate OS processes, each
with its own projector state on the GPU (scripts/solve.jl:121-148).
``PDFPBatch`` keeps the instance's scaled matrix in HBM once and runs the K
trajectories' PDHG iterations in the same fused launches (FP32 iterates,
FP64 reductions and polish).  Up to 8 trajectories share a batch of width 8
(a column's 8 values are 32 bytes), 9 to 16 one of width 16 (64 bytes, half
of the line a random gather fills).  Each trajectory has its own target, iterates,
primal weight, restart state and termination, and follows ``pdfp_solve!``
(src/pdfp.jl:252-288) on its own.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, dptr, f64, i64ptr, u8ptr
from .pdfp import SolverParameters, _csc_arrays


class PDFPBatch:
    """K ∈ 1..16 trajectories on the scaled data of one instance."""

    MAX = 16

    def __init__(self, K: int, A: sp.csc_matrix, b, is_eq, is_int, var_lb, var_ub, w, *,
                 device: int = 0):
        lib = _lib.load()
        self._lib = lib
        m, n = A.shape
        self.K, self.m, self.n = int(K), int(m), int(n)
        colptr, rowval, nzval = _csc_arrays(A)
        eq = np.ascontiguousarray(is_eq, dtype=np.uint8)
        ii = np.ascontiguousarray(is_int, dtype=np.uint8)
        ctx = C.c_void_p()
        st = lib.pdfp_batch_create(C.byref(ctx), self.K, m, n, len(nzval), i64ptr(colptr),
                                   i64ptr(rowval), dptr(nzval), 0, dptr(f64(b)), u8ptr(eq),
                                   u8ptr(ii), dptr(f64(var_lb)), dptr(f64(var_ub)), dptr(f64(w)),
                                   device)
        check(st, lib.pdfp_batch_last_error, None)
        self._ctx = ctx
        self.termination_reason = ["UNSPECIFIED"] * self.K

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdfp_batch_destroy(ctx)
            self._ctx = None

    def close(self):
        self.__del__()

    def _check(self, st):
        check(st, self._lib.pdfp_batch_last_error, self._ctx)

    def original_primal_residual(self, D_r) -> "PDFPBatch":
        d = f64(D_r)
        self._check(self._lib.pdfp_batch_set_original_primal_residual(self._ctx, dptr(d)))
        return self

    def set_x(self, k: int, x_scaled) -> None:
        x = f64(x_scaled)
        self._check(self._lib.pdfp_batch_set_x(self._ctx, k, dptr(x)))

    def set_target(self, k: int, x_tilde_scaled) -> None:
        x = f64(x_tilde_scaled)
        self._check(self._lib.pdfp_batch_set_target(self._ctx, k, dptr(x)))
        self.termination_reason[k] = "UNSPECIFIED"

    def set_column_scale(self, D_c) -> None:
        """The column scaling D_c (n positive values) for set_target_original."""
        d = f64(D_c)
        if d.shape != (self.n,):
            raise ValueError(f"D_c has shape {d.shape}, expected ({self.n},)")
        self._check(self._lib.pdfp_batch_set_column_scale(self._ctx, dptr(d)))

    def set_target_original(self, k: int, x_tilde) -> None:
        """Trajectory k's set_target!(projector, x̃ ./ D_c) (src/pump.jl:331)
        with x̃ in original units and the division on the device."""
        x = f64(x_tilde)
        if x.shape != (self.n,):
            raise ValueError(f"x_tilde has shape {x.shape}, expected ({self.n},)")
        self._check(self._lib.pdfp_batch_set_target_original(self._ctx, k, dptr(x)))
        self.termination_reason[k] = "UNSPECIFIED"

    def warm_start(self, k: int) -> None:
        self._check(self._lib.pdfp_batch_warm_start(self._ctx, k))
        self.termination_reason[k] = "UNSPECIFIED"

    def solve(self, p: SolverParameters, primal_weight, active=None) -> list:
        """pdfp_solve! for each active trajectory; returns their termination
        reasons (the inactive ones keep theirs)."""
        pw = np.broadcast_to(np.asarray(primal_weight, dtype=np.float64), (self.K,)).copy()
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        term = (C.c_int32 * self.K)()
        total = (C.c_int64 * self.K)()
        cp = p.to_c()
        self._check(self._lib.pdfp_batch_solve(self._ctx, C.byref(cp), dptr(pw),
                                               u8ptr(act) if act is not None else None, term,
                                               total))
        self.termination_reason = [_lib.TERMINATION[int(t)] for t in term]
        return list(self.termination_reason)

    def begin(self, k: int, p: SolverParameters, primal_weight: float) -> None:
        """Start trajectory k's solve with its own parameters (after
        set_target or warm_start); it runs in the next run()."""
        cp = p.to_c()
        self._check(self._lib.pdfp_batch_begin(self._ctx, k, C.byref(cp), float(primal_weight)))
        self.termination_reason[k] = "UNSPECIFIED"

    def run(self, until_any: bool = True, one_block: bool = False) -> tuple[list, int]:
        """Run the begun trajectories in lockstep blocks until one ends
        (until_any), for one block (one_block: a caller with trajectories waiting
        to begin polls between blocks) or until none runs.  Returns
        (termination reasons, with "UNSPECIFIED" for those still running;
        the number still running)."""
        mode = 2 if one_block else int(bool(until_any))
        term = (C.c_int32 * self.K)()
        total = (C.c_int64 * self.K)()
        left = C.c_int32()
        self._check(self._lib.pdfp_batch_run(self._ctx, mode, term, total,
                                             C.byref(left)))
        self.termination_reason = [_lib.TERMINATION[int(t)] for t in term]
        self.last_total_count = [int(t) for t in total]
        return list(self.termination_reason), int(left.value)

    def pdhg_x(self, k: int) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self._lib.pdfp_batch_get_x(self._ctx, k, dptr(out)))
        return out

    def vector(self, k: int, name: str) -> np.ndarray:
        ids = {"pdhg_x": 0, "pdhg_y": 1, "pdhg_z": 2, "current_x": 3, "current_y": 4,
               "current_z": 5, "dual_slack": 6, "x_tilde": 7}
        out = np.empty(self.m if name in ("pdhg_z", "current_z") else self.n)
        self._check(self._lib.pdfp_batch_get_vector(self._ctx, k, ids[name], dptr(out)))
        return out

    def info(self, k: int) -> _lib.PdfpInfo:
        out = _lib.PdfpInfo()
        self._check(self._lib.pdfp_batch_get_info(self._ctx, k, C.byref(out)))
        return out

    def total_count(self, k: int) -> int:
        return int(self.info(k).total_count)

    def set_kernel_path(self, path: str) -> None:
        """"auto" (the default): when the instance fits in LDS, each block
        runs as one launch with one workgroup per running trajectory;
        "general": always the batched passes."""
        codes = {"auto": 0, "general": 1}
        if path not in codes:
            raise ValueError("kernel path must be 'auto' or 'general'")
        self._check(self._lib.pdfp_batch_set_kernel_path(self._ctx, codes[path]))

    def set_block_graph(self, on: bool) -> None:
        """Replay each block of iterations (with its evaluations) as one
        captured HIP graph (the default), or launch every kernel on the
        stream.  Both give identical results."""
        self._check(self._lib.pdfp_batch_set_block_graph(self._ctx, int(bool(on))))

    def set_cooperative(self, wg_per_cu) -> None:
        """Run the middle iterations of each block of the batched passes in
        one cooperative launch of `wg_per_cu` workgroups per CU (True or -1:
        one fewer than fit), with a grid barrier between the passes, or launch
        each pass (False or 0, the default).  Both give identical results."""
        n = -1 if wg_per_cu is True else int(wg_per_cu)
        self._check(self._lib.pdfp_batch_set_cooperative(self._ctx, n))

    def set_column_slices(self, on) -> None:
        """Run A x̄ over A's 8 column slices, one per XCD (FP32 partial sums
        added in FP64), or in one pass (False).  256 or 512 also sets the
        nonzeros per workgroup of the slices' pass.  The autotuner picks the
        faster where x̄ outgrows an XCD's L2; the iterates differ in the last
        bits."""
        v = on if isinstance(on, int) and not isinstance(on, bool) else int(bool(on))
        self._check(self._lib.pdfp_batch_set_column_slices(self._ctx, v))

    def set_slice_min_len(self, min_len: int) -> None:
        """With column slices on, run only the rows of A with more than
        `min_len` nonzeros over the slices and the others in one pass (0:
        every row over the slices).  The autotuner picks it with the
        slices."""
        self._check(self._lib.pdfp_batch_set_slice_min_len(self._ctx, int(min_len)))

    def set_nontemporal(self, on: bool) -> None:
        """Load the passes' per-row and per-column operands with
        non-temporal loads (True) or plain loads (False); same results.  The
        autotuner picks."""
        self._check(self._lib.pdfp_batch_set_nontemporal(self._ctx, int(bool(on))))

    def set_schedule(self, npb_a: int, npb_at: int, persist: int = 0, lane_rows_a: int = 0,
                     lane_rows_at: int = 0) -> None:
        """Replace the autotuned schedule of the batched passes (tests and
        tuning): nonzeros per workgroup of the A and Aᵀ passes (256 or 512),
        persistent workgroups per CU, and the row-length bound of each pass's
        lane-per-(row, trajectory) reduction (0: off)."""
        self._check(self._lib.pdfp_batch_set_schedule(self._ctx, npb_a, npb_at, persist,
                                                      lane_rows_a, lane_rows_at))

    def time_kernels(self, iterations: int) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.pdfp_batch_time_kernels(self._ctx, iterations, C.byref(a),
                                                      C.byref(b)))
        return a.value, b.value

    def time_iterations(self, iterations: int) -> float:
        ms = C.c_double()
        self._check(self._lib.pdfp_batch_time_iterations(self._ctx, iterations, C.byref(ms)))
        return ms.value
