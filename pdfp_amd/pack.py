"""A pack of independent projector instances of uneven sizes on the device.

The reference solves independent instances in separate processes
(scripts/solve.jl:121-148).  ``PDFPPack`` keeps K scaled instances as one
block-diagonal matrix in HBM and runs all their PDHG iterations in the same
fused launches (FP32 iterates, FP64 reductions and polish).  Each instance
has its own target, iterates, step size, primal weight, restart state and
termination, and follows ``pdfp_solve!`` (src/pdfp.jl:252-288) on its own.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr, f64, u8ptr
from .pdfp import SolverParameters, _csc_arrays


class PDFPPack:
    """K projector instances, each given as the arguments of ``PDFPState``
    (a ``ProjectorProblem``: A, b, is_eq, is_int, var_lb, var_ub, w)."""

    def __init__(self, problems, *, device: int = 0):
        lib = _lib.load()
        self._lib = lib
        self.K = len(problems)
        self.m = [int(p.A.shape[0]) for p in problems]
        self.n = [int(p.A.shape[1]) for p in problems]
        # keep every array alive for the call
        self._keep = []
        arrs = {k: [] for k in ("colptr", "rowval", "nzval", "b", "eq", "ii", "lb", "ub", "w")}
        for p in problems:
            colptr, rowval, nzval = _csc_arrays(p.A)
            arrs["colptr"].append(colptr)
            arrs["rowval"].append(rowval)
            arrs["nzval"].append(nzval)
            arrs["b"].append(f64(p.b))
            arrs["eq"].append(np.ascontiguousarray(p.is_eq, dtype=np.uint8))
            arrs["ii"].append(np.ascontiguousarray(p.is_int, dtype=np.uint8))
            arrs["lb"].append(f64(p.var_lb))
            arrs["ub"].append(f64(p.var_ub))
            arrs["w"].append(f64(p.w))
        self._keep.append(arrs)

        def ptrs(name, ctype):
            lst = arrs[name]
            out = (C.POINTER(ctype) * self.K)()
            for k, a in enumerate(lst):
                out[k] = a.ctypes.data_as(C.POINTER(ctype))
            return out

        m = np.asarray(self.m, dtype=np.int64)
        n = np.asarray(self.n, dtype=np.int64)
        self._keep.append((m, n))
        ctx = C.c_void_p()
        st = lib.pdfp_pack_create(C.byref(ctx), self.K, m.ctypes.data_as(C.POINTER(C.c_int64)),
                                  n.ctypes.data_as(C.POINTER(C.c_int64)),
                                  ptrs("colptr", C.c_int64), ptrs("rowval", C.c_int64),
                                  ptrs("nzval", C.c_double), 0, ptrs("b", C.c_double),
                                  ptrs("eq", C.c_uint8), ptrs("ii", C.c_uint8),
                                  ptrs("lb", C.c_double), ptrs("ub", C.c_double),
                                  ptrs("w", C.c_double), device)
        check(st, lib.pdfp_pack_last_error, None)
        self._keep = None  # the library copied what it needs
        self._ctx = ctx
        self.termination_reason = ["UNSPECIFIED"] * self.K

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdfp_pack_destroy(ctx)
            self._ctx = None

    def close(self):
        self.__del__()

    def _check(self, st):
        check(st, self._lib.pdfp_pack_last_error, self._ctx)

    def original_primal_residual(self, k: int, D_r) -> "PDFPPack":
        d = f64(D_r)
        self._check(self._lib.pdfp_pack_set_original_primal_residual(self._ctx, k, dptr(d)))
        return self

    def set_x(self, k: int, x_scaled) -> None:
        x = f64(x_scaled)
        self._check(self._lib.pdfp_pack_set_x(self._ctx, k, dptr(x)))

    def set_target(self, k: int, x_tilde_scaled) -> None:
        x = f64(x_tilde_scaled)
        self._check(self._lib.pdfp_pack_set_target(self._ctx, k, dptr(x)))
        self.termination_reason[k] = "UNSPECIFIED"

    def warm_start(self, k: int) -> None:
        self._check(self._lib.pdfp_pack_warm_start(self._ctx, k))
        self.termination_reason[k] = "UNSPECIFIED"

    def solve(self, p: SolverParameters, primal_weight, active=None) -> list:
        """pdfp_solve! for each active instance; returns their termination
        reasons (the inactive ones keep theirs)."""
        pw = np.broadcast_to(np.asarray(primal_weight, dtype=np.float64), (self.K,)).copy()
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        term = (C.c_int32 * self.K)()
        total = (C.c_int64 * self.K)()
        cp = p.to_c()
        self._check(self._lib.pdfp_pack_solve(self._ctx, C.byref(cp), dptr(pw),
                                              u8ptr(act) if act is not None else None, term,
                                              total))
        self.termination_reason = [_lib.TERMINATION[int(t)] for t in term]
        self.last_total_count = [int(t) for t in total]
        return list(self.termination_reason)

    def begin(self, k: int, p: SolverParameters, primal_weight: float) -> None:
        """Start instance k's solve with its own parameters (after set_target
        or warm_start); it runs in the next run()."""
        cp = p.to_c()
        self._check(self._lib.pdfp_pack_begin(self._ctx, k, C.byref(cp), float(primal_weight)))
        self.termination_reason[k] = "UNSPECIFIED"

    def run(self, until_any: bool = True, one_block: bool = False) -> tuple[list, int]:
        """Run the begun instances in lockstep blocks until one ends
        (until_any), for one block (one_block: a caller with instances waiting
        to begin polls between blocks) or until none runs.  Returns
        (termination reasons, with "UNSPECIFIED" for those still running;
        the number still running)."""
        mode = 2 if one_block else int(bool(until_any))
        term = (C.c_int32 * self.K)()
        total = (C.c_int64 * self.K)()
        left = C.c_int32()
        self._check(self._lib.pdfp_pack_run(self._ctx, mode, term, total,
                                            C.byref(left)))
        self.termination_reason = [_lib.TERMINATION[int(t)] for t in term]
        self.last_total_count = [int(t) for t in total]
        return list(self.termination_reason), int(left.value)

    def set_kernel_path(self, path: str) -> None:
        """"auto" (the default): when every instance fits in LDS, each block
        runs as one launch with one workgroup per running instance;
        "general": always the general passes."""
        codes = {"auto": 0, "general": 1}
        if path not in codes:
            raise ValueError("kernel path must be 'auto' or 'general'")
        self._check(self._lib.pdfp_pack_set_kernel_path(self._ctx, codes[path]))

    def set_block_graph(self, on: bool) -> None:
        """Replay each block of iterations (with its evaluations) as one
        captured HIP graph (the default), or launch every kernel on the
        stream.  Both give identical results."""
        self._check(self._lib.pdfp_pack_set_block_graph(self._ctx, int(bool(on))))

    def pdhg_x(self, k: int) -> np.ndarray:
        out = np.empty(self.n[k])
        self._check(self._lib.pdfp_pack_get_x(self._ctx, k, dptr(out)))
        return out

    def vector(self, k: int, name: str) -> np.ndarray:
        ids = {"pdhg_x": 0, "pdhg_y": 1, "pdhg_z": 2, "current_x": 3, "current_y": 4,
               "current_z": 5, "dual_slack": 6, "x_tilde": 7}
        out = np.empty(self.m[k] if name in ("pdhg_z", "current_z") else self.n[k])
        self._check(self._lib.pdfp_pack_get_vector(self._ctx, k, ids[name], dptr(out)))
        return out

    def info(self, k: int) -> _lib.PdfpInfo:
        out = _lib.PdfpInfo()
        self._check(self._lib.pdfp_pack_get_info(self._ctx, k, C.byref(out)))
        return out

    def total_count(self, k: int) -> int:
        return int(self.info(k).total_count)

    def time_iterations(self, iterations: int) -> float:
        ms = C.c_double()
        self._check(self._lib.pdfp_pack_time_iterations(self._ctx, iterations, C.byref(ms)))
        return ms.value

    def time_kernels(self, iterations: int) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.pdfp_pack_time_kernels(self._ctx, iterations, C.byref(a),
                                                     C.byref(b)))
        return a.value, b.value
