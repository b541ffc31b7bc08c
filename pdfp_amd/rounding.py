"""The pump's rounders, on the host, with src/rounding.jl's arithmetic.

Sequential rounding orders the integer columns (binary columns ahead of the
others, then by distance of the LP value from its nearest integer) and
settles them in that order: each column is set to its rounded LP value and
the bound propagation runs before the next one is settled.  The guarded
variant settles, in a single propagation, all columns whose LP value is an
exact integer, and only then walks the remaining columns one by one; should
that first propagation prove some row infeasible, everything it did is
thrown away and the sequential rounder runs from scratch instead.

Random numbers: the reference draws from Julia's MersenneTwister; here the
caller's ``numpy.random.Generator`` is used at the same points
(``rand(rng, 1:len)`` ↔ ``rng.integers(1, len + 1)``).  Every rounding that
draws nothing (iteration < swap_iteration, no flips or perturbations) is
therefore identical to the reference's.

``Rounder(backend="native")`` (the default) runs the propagation and the
rounding passes in C++ (``pdfp_rounder_*`` in libpdfp_hip); the ranking, its
random draws and the batch policy's timing stay here.  ``backend="python"``
runs everything in this module (PropagationState); the two are bit-identical.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .milp import MILP
from .propagation import BOUND_TOLERANCE, PropagationData, PropagationState


def jround(v: float) -> float:
    """Julia's round (ties to even, the sign of zero kept)."""
    return float(np.rint(v))


def jclamp(x: float, lo: float, hi: float) -> float:
    return hi if x > hi else (lo if x < lo else x)


class BatchPolicy:
    """Chooses at each call between trying the all-at-once step of the
    guarded rounder and going straight to sequential rounding, from running
    averages of what each choice cost before (src/rounding.jl:22-63)."""

    def __init__(self, xi: float = 0.3, nu: float = 0.9, reprobe: int = 20):
        self.xi, self.nu, self.reprobe = xi, nu, reprobe
        self.n_failed = 0.0
        self.n_held = 0.0
        self.s = math.nan
        self.h = math.nan
        self.w = math.nan
        self.skipped = 0

    def _ema(self, old: float, x: float) -> float:
        return x if math.isnan(old) else (1.0 - self.xi) * old + self.xi * x

    def wants_batch(self) -> bool:
        if self.skipped >= self.reprobe:
            return True
        if math.isnan(self.s) or self.n_failed == 0.0:
            return True
        h = 0.0 if math.isnan(self.h) else self.h
        p = self.n_failed / (self.n_failed + self.n_held)
        return p * (self.w + self.s) + (1.0 - p) * h < self.s

    def observe_batch(self, failed: bool, cost: float, wasted: float) -> None:
        self.n_failed *= self.nu
        self.n_held *= self.nu
        self.skipped = 0
        if failed:
            self.n_failed += 1.0
            self.w = self._ema(self.w, wasted)
            self.s = self._ema(self.s, max(cost - wasted, 0.0))
        else:
            self.n_held += 1.0
            self.h = self._ema(self.h, cost)

    def observe_skip(self, cost: float) -> None:
        self.s = self._ema(self.s, cost)
        self.skipped += 1


class Rounder:
    """One pump run's rounder: the model's propagation data, the schedule,
    the optional violation-guided columns and the counters live here from
    call to call (src/rounding.jl:78-98)."""

    def __init__(self, milp: MILP, *, schedule: str = "all", max_pops: int = 2**63 - 1,
                 swap_iteration: int = 10, policy: BatchPolicy | None | str = "default",
                 backend: str = "native"):
        if backend not in ("native", "python"):
            raise ValueError("backend must be 'native' or 'python'")
        self.milp = milp
        self.data: PropagationData | None = None
        self.schedule = schedule                # "all" or "regroup"
        self._direction: list | None = None     # violation-guided columns, or None
        self.max_pops = max_pops
        self.swap_iteration = swap_iteration
        self.policy = BatchPolicy() if policy == "default" else policy
        self.backend = backend
        self.n_calls = 0
        self.n_batched = 0
        self.n_fallbacks = 0
        self._int_idx = milp.int_idx.tolist()
        self._int_arr = np.asarray(milp.int_idx, dtype=np.int64)
        self._binary = [(milp.l[j] == 0.0 and milp.u[j] == 1.0) for j in range(len(milp.c))]
        self._bin_penalty = np.where((milp.l == 0.0) & (milp.u == 1.0), 10.0, 0.0)[self._int_arr]
        self._b = milp.b.tolist()
        self._is_le = [s == "le" for s in milp.row_sense.tolist()]
        self._native = None
        if backend == "native":
            self._native = _NativeRounder(milp, max_pops)

    @property
    def direction(self) -> list | None:
        return self._direction

    @direction.setter
    def direction(self, value) -> None:
        self._direction = value
        if self._native is not None:
            self._native.set_direction(value)

    def _new_state(self) -> PropagationState:
        if self.data is None:
            self.data = PropagationData(self.milp)
        return PropagationState(self.data, self.max_pops)

    def ranking(self, x_star, rng, iteration: int) -> np.ndarray:
        """Ascending fractionality, binaries first (a stable sort, as
        sortperm); from swap_iteration on, |I|/10 random transpositions
        (src/rounding.jl:105-115)."""
        v = np.asarray(x_star, dtype=np.float64)[self._int_arr]
        keys = np.abs(v - np.rint(v)) - self._bin_penalty
        ranked = self._int_arr[np.argsort(keys, kind="stable")]
        n = len(ranked)
        if iteration >= self.swap_iteration and n >= 10:
            # The draws a₁, b₁, a₂, b₂, … in one call: a Generator's vector
            # of bounded integers is the sequence of its scalar draws.
            ab = (rng.integers(1, n + 1, size=2 * max(1, n // 10)) - 1).tolist()
            r = memoryview(ranked)          # swaps in place, one at a time
            for a, b in zip(ab[0::2], ab[1::2]):
                r[a], r[b] = r[b], r[a]
        return ranked

    def _row_violation(self, i: int, activity: float) -> float:
        return max(0.0, activity - self._b[i]) if self._is_le[i] else abs(activity - self._b[i])

    def _guided_value(self, s: PropagationState, j: int, target: float, activity) -> float:
        """src/rounding.jl:123-141."""
        if not s.u[j] - s.l[j] > BOUND_TOLERANCE:
            return target
        lo = float(math.ceil(s.l[j] - BOUND_TOLERANCE))
        hi = float(math.floor(s.u[j] + BOUND_TOLERANCE))
        if not hi > lo:
            return target
        d = self.data
        rows, vals = d.col_rows[j], d.col_vals[j]

        def score(v):
            t = 0.0
            for i, a in zip(rows, vals):
                t += self._row_violation(i, activity[i] + a * v) - self._row_violation(i, activity[i])
            return t

        score_lo = score(lo)
        score_hi = score(hi)
        if score_lo < score_hi - 1e-12:
            return lo
        if score_hi < score_lo - 1e-12:
            return hi
        return target

    def _add_activity(self, activity, j: int, v: float) -> None:
        d = self.data
        for i, a in zip(d.col_rows[j], d.col_vals[j]):
            activity[i] += a * v

    def _round_sequentially(self, s: PropagationState, xt: list, idx, x_star, activity) -> None:
        """src/rounding.jl:152-174."""
        for j in idx:
            if s.is_fixed(j):
                xt[j] = s.l[j]
                continue
            target = jround(jclamp(x_star[j], s.l[j], s.u[j]))
            if target < s.l[j] - BOUND_TOLERANCE or target > s.u[j] + BOUND_TOLERANCE:
                other = jround(jclamp(x_star[j] + (1.0 if target < s.l[j] else -1.0),
                                      s.l[j], s.u[j]))
                target = jclamp(other, s.l[j], s.u[j])
            if activity is not None and self.direction[j]:
                target = self._guided_value(s, j, target, activity)
            s.fix(j, target)
            s.propagate()
            if activity is not None:
                self._add_activity(activity, j, target)
            xt[j] = target

    def _read_integers(self, s: PropagationState, xt: list, x_star) -> list:
        for j in self._int_idx:
            xt[j] = s.l[j] if s.is_fixed(j) else jround(jclamp(x_star[j], s.l[j], s.u[j]))
        return xt

    def _activity(self):
        return None if self.direction is None else [0.0] * len(self.milp.b)

    def round_sequential(self, x_star, rng, iteration: int) -> np.ndarray:
        """FP 2.0's CPR of x_star (src/rounding.jl:191-197)."""
        ranked = self.ranking(x_star, rng, iteration)
        if self._native is not None:
            return self._native.round_sequential(x_star, ranked)
        s = self._new_state()
        xs = [float(v) for v in x_star]
        xt = list(xs)
        self._round_sequentially(s, xt, ranked.tolist(), xs, self._activity())
        return np.array(self._read_integers(s, xt, xs))

    def round_guarded(self, x_star, rng, iteration: int) -> np.ndarray:
        """Guarded two-phase CPR (src/rounding.jl:205-256)."""
        self.n_calls += 1
        P = self.policy
        t0 = time.perf_counter_ns() if P is not None else 0
        if P is not None and not P.wants_batch():
            xt = self.round_sequential(x_star, rng, iteration)
            P.observe_skip((time.perf_counter_ns() - t0) / 1e9)
            return xt
        ranked_arr = self.ranking(x_star, rng, iteration)
        if self._native is not None:
            xt_native, infeasible = self._native.round_guarded(x_star, ranked_arr,
                                                               self.schedule == "regroup")
            if infeasible:               # a row went infeasible: drop the all-at-once step
                self.n_fallbacks += 1
                wasted = 0.0 if P is None else (time.perf_counter_ns() - t0) / 1e9
                out = self.round_sequential(x_star, rng, iteration)
                if P is not None:
                    P.observe_batch(True, (time.perf_counter_ns() - t0) / 1e9, wasted)
                return out
            self.n_batched += 1
            if P is not None:
                P.observe_batch(False, (time.perf_counter_ns() - t0) / 1e9, 0.0)
            return xt_native
        s = self._new_state()
        xs = [float(v) for v in x_star]
        xt = list(xs)
        activity = self._activity()
        ranked = ranked_arr.tolist()

        def integral(j):
            return abs(xs[j] - jround(xs[j])) == 0.0

        if self.schedule == "regroup":
            last = 0
            for k, j in enumerate(ranked):
                if integral(j):
                    last = k + 1
            before = [j for j in ranked[:last] if not integral(j)]
            batch = [j for j in ranked[:last] if integral(j)]
            after = ranked[last:]
        else:
            before = []
            batch = [j for j in ranked if integral(j)]
            after = [j for j in ranked if not integral(j)]

        self._round_sequentially(s, xt, before, xs, activity)
        for j in batch:
            if s.is_fixed(j):
                xt[j] = s.l[j]
                continue
            target = jround(jclamp(xs[j], s.l[j], s.u[j]))
            s.fix(j, target)
            if activity is not None:
                self._add_activity(activity, j, target)
            xt[j] = target
        s.propagate()

        if s.infeasible:                 # a row went infeasible: drop the all-at-once step
            self.n_fallbacks += 1
            wasted = 0.0 if P is None else (time.perf_counter_ns() - t0) / 1e9
            out = self.round_sequential(x_star, rng, iteration)
            if P is not None:
                P.observe_batch(True, (time.perf_counter_ns() - t0) / 1e9, wasted)
            return out
        self.n_batched += 1
        self._round_sequentially(s, xt, after, xs, activity)
        self._read_integers(s, xt, xs)
        if P is not None:
            P.observe_batch(False, (time.perf_counter_ns() - t0) / 1e9, 0.0)
        return np.array(xt)


class _NativeRounder:
    """The C++ rounder of libpdfp_hip (pdfp_rounder_*)."""

    def __init__(self, milp: MILP, max_pops: int):
        import ctypes as C

        from . import _lib

        self._C = C
        self._L = _lib
        lib = _lib.load()
        self._lib = lib
        A = milp.A.tocsc()
        self.n = A.shape[1]
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
        nzval = np.ascontiguousarray(A.data, dtype=np.float64)
        b = np.ascontiguousarray(milp.b, dtype=np.float64)
        is_le = np.ascontiguousarray(milp.row_sense == "le", dtype=np.uint8)
        lb = np.ascontiguousarray(milp.l, dtype=np.float64)
        ub = np.ascontiguousarray(milp.u, dtype=np.float64)
        is_int = np.zeros(self.n, dtype=np.uint8)
        is_int[milp.int_idx] = 1
        ctx = C.c_void_p()
        st = lib.pdfp_rounder_create(C.byref(ctx), A.shape[0], self.n, _lib.i64ptr(colptr),
                                     _lib.i64ptr(rowval), _lib.dptr(nzval), 0, _lib.dptr(b),
                                     _lib.u8ptr(is_le), _lib.dptr(lb), _lib.dptr(ub),
                                     _lib.u8ptr(is_int), int(min(max_pops, 2**63 - 1)))
        _lib.check(st, lib.pdfp_rounder_last_error, None)
        self._ctx = ctx

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdfp_rounder_destroy(ctx)
            self._ctx = None

    def _check(self, st):
        self._L.check(st, self._lib.pdfp_rounder_last_error, self._ctx)

    def set_direction(self, direction) -> None:
        if direction is None:
            self._check(self._lib.pdfp_rounder_set_direction(self._ctx, None))
        else:
            d = np.ascontiguousarray(direction, dtype=np.uint8)
            self._check(self._lib.pdfp_rounder_set_direction(self._ctx, self._L.u8ptr(d)))

    def round_sequential(self, x_star, ranked) -> np.ndarray:
        xs = np.ascontiguousarray(x_star, dtype=np.float64)
        rk = np.ascontiguousarray(ranked, dtype=np.int64)
        xt = np.empty(self.n)
        self._check(self._lib.pdfp_round_sequential(self._ctx, self._L.dptr(xs),
                                                    self._L.i64ptr(rk), len(rk),
                                                    self._L.dptr(xt)))
        return xt

    def round_guarded(self, x_star, ranked, regroup: bool):
        C = self._C
        xs = np.ascontiguousarray(x_star, dtype=np.float64)
        rk = np.ascontiguousarray(ranked, dtype=np.int64)
        xt = np.empty(self.n)
        bad = C.c_int32(0)
        self._check(self._lib.pdfp_round_guarded(self._ctx, self._L.dptr(xs),
                                                 self._L.i64ptr(rk), len(rk), int(regroup),
                                                 self._L.dptr(xt), C.byref(bad)))
        return xt, bool(bad.value)
