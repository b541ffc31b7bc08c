"""Domain propagation on the rows of A, as the rounders use it (host, FP64,
src/propagation.jl's arithmetic).

For every row the state tracks the smallest and largest value its left-hand
side can take over the current variable domains.  Tightening a domain
adjusts those two numbers in place and appends the affected rows, first in
first out, to the queue of rows to re-examine.  Rows whose coefficients are
all 1 on binary columns, and ≤ rows with positive coefficients on binary
columns, are examined by shortcut routines that derive exactly the bounds
the general routine would.  The data are plain Python lists because every
loop visits a single row or column.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np
import scipy.sparse as sp

from .milp import MILP

BOUND_TOLERANCE = 1e-6
INFEASIBILITY_TOLERANCE = 1e-6
MAX_TIGHTENINGS = 10       # after this many tightenings of one bound, only a fixing re-queues its rows
DOMAIN_INFINITY = 1e12     # finite proxy for ±inf bounds inside activity sums

ROW_GENERAL = 0
ROW_CARDINALITY = 1
ROW_KNAPSACK = 2


class PropagationData:
    """The model-derived, call-invariant half of the propagation: row and
    column nonzero lists, row kinds and the starting domains
    (src/propagation.jl:19-84)."""

    def __init__(self, milp: MILP):
        n = len(milp.c)
        m = len(milp.b)
        self.n, self.m = n, m
        self.l = [v if math.isfinite(v) else -DOMAIN_INFINITY for v in milp.l.tolist()]
        self.u = [v if math.isfinite(v) else DOMAIN_INFINITY for v in milp.u.tolist()]
        is_int = [False] * n
        for j in milp.int_idx.tolist():
            is_int[j] = True
        self.is_int = is_int
        self.b = milp.b.tolist()
        self.is_le = [s == "le" for s in milp.row_sense.tolist()]
        A = milp.A
        # columns of A (CSC) and rows of A (CSR of Aᵀ with ascending columns)
        self.col_rows = []
        self.col_vals = []
        ip, ix, dv = A.indptr.tolist(), A.indices.tolist(), A.data.tolist()
        for j in range(n):
            self.col_rows.append(ix[ip[j]:ip[j + 1]])
            self.col_vals.append(dv[ip[j]:ip[j + 1]])
        At = sp.csr_matrix(A)
        At.sort_indices()
        rp, rc, rv = At.indptr.tolist(), At.indices.tolist(), At.data.tolist()
        self.row_cols = [rc[rp[i]:rp[i + 1]] for i in range(m)]
        self.row_vals = [rv[rp[i]:rp[i + 1]] for i in range(m)]
        l, u = self.l, self.u
        L_min = [0.0] * m
        L_max = [0.0] * m
        row_kind = [ROW_GENERAL] * m
        row_maxcoef = [0.0] * m
        row_maxroom = [0.0] * m
        ml, mu = milp.l.tolist(), milp.u.tolist()
        for i in range(m):
            lmin = 0.0
            lmax = 0.0
            cols, vals = self.row_cols[i], self.row_vals[i]
            for j, a in zip(cols, vals):
                if a > 0:
                    lmin += a * l[j]
                    lmax += a * u[j]
                else:
                    lmin += a * u[j]
                    lmax += a * l[j]
            L_min[i] = lmin
            L_max[i] = lmax
            all_one = all_pos = all_bin = True
            maxcoef = 0.0
            for j, a in zip(cols, vals):
                if abs(a) > maxcoef:
                    maxcoef = abs(a)
                if a != 1.0:
                    all_one = False
                if not a > 0.0:
                    all_pos = False
                if not (is_int[j] and ml[j] == 0.0 and mu[j] == 1.0):
                    all_bin = False
                if not (all_bin and (all_one or all_pos)):
                    break
            row_maxcoef[i] = maxcoef
            if all_bin and all_one:
                row_kind[i] = ROW_CARDINALITY
            elif all_bin and all_pos and self.is_le[i]:
                row_kind[i] = ROW_KNAPSACK
            room = 0.0
            for j, a in zip(cols, vals):
                width = mu[j] - ml[j]
                r = abs(a) * width if math.isfinite(width) else math.inf
                if r > room:
                    room = r
            row_maxroom[i] = room
        self.L_min, self.L_max = L_min, L_max
        self.row_kind, self.row_maxcoef, self.row_maxroom = row_kind, row_maxcoef, row_maxroom


class PropagationState:
    """Per-call state: the current lower/upper bounds, the row activity range and the FIFO of rows to revisit
    (src/propagation.jl:87-107)."""

    def __init__(self, d: PropagationData, max_pops: int):
        self.data = d
        self.l = list(d.l)
        self.u = list(d.u)
        self.L_min = list(d.L_min)
        self.L_max = list(d.L_max)
        self.in_queue = [False] * d.m
        self.queue: deque = deque()
        self.tighten_lo = [0] * d.n
        self.tighten_hi = [0] * d.n
        self.row_maxroom = list(d.row_maxroom)
        self.max_pops = max_pops
        self.infeasible = False

    def _enqueue(self, i: int) -> None:
        if not self.in_queue[i]:
            self.queue.append(i)
            self.in_queue[i] = True

    def _tighten_upper(self, j: int, new_u: float) -> bool:
        """src/propagation.jl:117-136."""
        u_old = self.u[j]
        if not new_u < u_old - BOUND_TOLERANCE:
            return False
        nu = math.floor(new_u + BOUND_TOLERANCE) if self.data.is_int[j] else new_u
        if nu < self.l[j] - BOUND_TOLERANCE:
            nu = self.l[j]
        nu = max(nu, self.l[j])
        if nu >= u_old - BOUND_TOLERANCE:
            return False
        nu = float(nu)
        self.u[j] = nu
        self.tighten_hi[j] += 1
        wake = self.tighten_hi[j] <= MAX_TIGHTENINGS or abs(nu - self.l[j]) < BOUND_TOLERANCE
        du = nu - u_old
        L_min, L_max = self.L_min, self.L_max
        for i, a in zip(self.data.col_rows[j], self.data.col_vals[j]):
            if a > 0:
                L_max[i] += a * du
            else:
                L_min[i] += a * du
            if wake:
                self._enqueue(i)
        return True

    def _tighten_lower(self, j: int, new_l: float) -> bool:
        """src/propagation.jl:138-156."""
        l_old = self.l[j]
        if not new_l > l_old + BOUND_TOLERANCE:
            return False
        nl = math.ceil(new_l - BOUND_TOLERANCE) if self.data.is_int[j] else new_l
        if nl > self.u[j] + BOUND_TOLERANCE:
            nl = self.u[j]
        nl = min(nl, self.u[j])
        if nl <= l_old + BOUND_TOLERANCE:
            return False
        nl = float(nl)
        self.l[j] = nl
        self.tighten_lo[j] += 1
        wake = self.tighten_lo[j] <= MAX_TIGHTENINGS or abs(self.u[j] - nl) < BOUND_TOLERANCE
        dl = nl - l_old
        L_min, L_max = self.L_min, self.L_max
        for i, a in zip(self.data.col_rows[j], self.data.col_vals[j]):
            if a > 0:
                L_min[i] += a * dl
            else:
                L_max[i] += a * dl
            if wake:
                self._enqueue(i)
        return True

    def _propagate_cardinality_row(self, i: int) -> None:
        """src/propagation.jl:160-177."""
        d = self.data
        b = d.b[i]
        slack_up = b - self.L_min[i]
        slack_dn = self.L_max[i] - b
        is_eq = not d.is_le[i]
        if slack_up < -INFEASIBILITY_TOLERANCE or (is_eq and slack_dn < -INFEASIBILITY_TOLERANCE):
            self.infeasible = True
            return
        up = slack_up < 1.0 - BOUND_TOLERANCE
        dn = is_eq and slack_dn < 1.0 - BOUND_TOLERANCE
        if not (up or dn):
            return
        for j in d.row_cols[i]:
            if up:
                self._tighten_upper(j, self.l[j] + slack_up)
            if dn:
                self._tighten_lower(j, self.u[j] - slack_dn)

    def _propagate_knapsack_row(self, i: int) -> None:
        """src/propagation.jl:180-192."""
        d = self.data
        slack_up = d.b[i] - self.L_min[i]
        if slack_up < -INFEASIBILITY_TOLERANCE:
            self.infeasible = True
            return
        if d.row_maxcoef[i] <= slack_up + BOUND_TOLERANCE:
            return
        for j, a in zip(d.row_cols[i], d.row_vals[i]):
            if a > slack_up + BOUND_TOLERANCE:
                self._tighten_upper(j, self.l[j])

    def _propagate_row(self, i: int) -> None:
        """src/propagation.jl:194-232."""
        d = self.data
        kind = d.row_kind[i]
        if kind == ROW_CARDINALITY:
            return self._propagate_cardinality_row(i)
        if kind == ROW_KNAPSACK:
            return self._propagate_knapsack_row(i)
        is_le = d.is_le[i]
        b = d.b[i]
        slack_up = b - self.L_min[i]
        slack_dn = self.L_max[i] - b
        if slack_up < -INFEASIBILITY_TOLERANCE or ((not is_le) and slack_dn < -INFEASIBILITY_TOLERANCE):
            self.infeasible = True
            return
        threshold = (slack_up if is_le else min(slack_up, slack_dn)) + BOUND_TOLERANCE
        if self.row_maxroom[i] <= threshold:
            return
        cols, vals = d.row_cols[i], d.row_vals[i]
        l, u = self.l, self.u
        max_room = 0.0
        for j, a in zip(cols, vals):
            room = abs(a) * (u[j] - l[j])
            if room > max_room:
                max_room = room
        self.row_maxroom[i] = max_room
        if is_le and max_room <= slack_up + BOUND_TOLERANCE:
            return
        for j, a in zip(cols, vals):
            if a > 0:
                self._tighten_upper(j, l[j] + slack_up / a)
            else:
                self._tighten_lower(j, u[j] + slack_up / a)
            if not is_le:
                if a > 0:
                    self._tighten_lower(j, u[j] - slack_dn / a)
                else:
                    self._tighten_upper(j, l[j] - slack_dn / a)

    def propagate(self) -> int:
        """Revisit queued rows, first in first out, until none is left or the pop budget is
        spent (src/propagation.jl:235-244)."""
        pops = 0
        q = self.queue
        while q and pops < self.max_pops:
            i = q.popleft()
            self.in_queue[i] = False
            self._propagate_row(i)
            pops += 1
        return pops

    def fix(self, j: int, value: float) -> bool:
        """Fix x_j = value and queue its rows (src/propagation.jl:252-274)."""
        l_old, u_old = self.l[j], self.u[j]
        if value < l_old - BOUND_TOLERANCE or value > u_old + BOUND_TOLERANCE:
            self.infeasible = True
            return False
        v = u_old if value > u_old else (l_old if value < l_old else value)
        if abs(v - l_old) < BOUND_TOLERANCE and abs(v - u_old) < BOUND_TOLERANCE:
            return True
        self.l[j] = v
        self.u[j] = v
        self.tighten_lo[j] += 1
        self.tighten_hi[j] += 1
        dl = v - l_old
        du = v - u_old
        L_min, L_max = self.L_min, self.L_max
        for i, a in zip(self.data.col_rows[j], self.data.col_vals[j]):
            if a > 0:
                L_min[i] += a * dl
                L_max[i] += a * du
            else:
                L_min[i] += a * du
                L_max[i] += a * dl
            self._enqueue(i)
        return True

    def is_fixed(self, j: int) -> bool:
        return abs(self.u[j] - self.l[j]) < BOUND_TOLERANCE


def activities(d: PropagationData, l, u):
    """(L_min, L_max) recomputed from scratch for the domains (l, u); the
    incremental values of PropagationState must agree (used by the tests)."""
    L_min = np.zeros(d.m)
    L_max = np.zeros(d.m)
    for i in range(d.m):
        lmin = lmax = 0.0
        for j, a in zip(d.row_cols[i], d.row_vals[i]):
            if a > 0:
                lmin += a * l[j]
                lmax += a * u[j]
            else:
                lmin += a * u[j]
                lmax += a * l[j]
        L_min[i], L_max[i] = lmin, lmax
    return L_min, L_max
