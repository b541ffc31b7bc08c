"""Synthetic MILP workloads of the projector benchmarks and parity tests.

Every instance is drawn from numpy's PCG64 generator seeded with the given
seed ("seed s" means this generator seeded with s).  The families:

- ``set_cover``         binary set covering: each column covers 3-15 distinct
                        rows chosen uniformly, A ∈ {0,1}, A x ≥ 1 written as
                        -A x ≤ -1, costs uniform integers in [1, 100].
                        (133 × 201 mirrors p0201's shape.)
- ``mixed``             rows of 10-30 distinct columns (or power-law lengths),
                        values N(0,1), a seeded 70 % of the columns binary and
                        30 % continuous in [0, 10], b = A x0 + |N(0,1)| for a
                        random point x0, so the LP is feasible.
- ``general_integer``   integer columns in [0, U_j], U_j uniform in 1-20,
                        integer coefficients in [-9, 9]∖{0}, 40 % equality rows
                        A_eq x = A_eq x0, the rest ≤ with |N(0,1)| slack.
- ``uneven_set_covers`` K set-cover instances of log-uniform sizes.
- ``binary_targets``    the pump-like targets x̃: uniform 0/1 on the binary
                        columns (rounded uniform integers on general ones).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .milp import MILP


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _distinct_columns(rng, lengths: np.ndarray, n: int) -> tuple[np.ndarray, np.ndarray]:
    """For each row i, ``lengths[i]`` distinct column indices in [0, n),
    uniform.  Returns (row_ptr, cols) with the columns of each row sorted."""
    lengths = np.asarray(lengths, dtype=np.int64)
    if np.any(lengths > n):
        raise ValueError("a row is longer than the number of columns")
    m = len(lengths)
    ptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lengths, out=ptr[1:])
    nnz = int(ptr[-1])
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=nnz, dtype=np.int64)
    while True:
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup_sorted = np.zeros(nnz, dtype=bool)
        dup_sorted[1:] = ks[1:] == ks[:-1]
        if not dup_sorted.any():
            cols = cols[order]
            break
        dup = np.zeros(nnz, dtype=bool)
        dup[order[dup_sorted]] = True
        cols[dup] = rng.integers(0, n, size=int(dup.sum()), dtype=np.int64)
    return ptr, cols


def _csr_to_csc(m: int, n: int, ptr, cols, vals) -> sp.csc_matrix:
    A = sp.csr_matrix((vals, cols, ptr), shape=(m, n)).tocsc()
    A.sort_indices()
    return A


def set_cover(m: int = 133, n: int = 201, seed: int = 1, min_cover: int = 3,
              max_cover: int = 15) -> MILP:
    """Binary set covering, -A x ≤ -1 (each row must be covered)."""
    rng = _rng(seed)
    k = rng.integers(min_cover, max_cover + 1, size=n)
    k = np.minimum(k, m)
    # column j covers k_j distinct rows: draw as "rows of Aᵀ"
    ptr, rows_of_col = _distinct_columns(rng, k, m)
    cover = np.zeros(m, dtype=np.int64)
    np.add.at(cover, rows_of_col, 1)
    extra_r, extra_c = [], []
    for i in np.flatnonzero(cover == 0):      # repair: add row i to a random column
        extra_r.append(int(i))
        extra_c.append(int(rng.integers(0, n)))
    cols_of_entry = np.repeat(np.arange(n, dtype=np.int64), k)
    I = np.concatenate([rows_of_col, np.array(extra_r, dtype=np.int64)])
    J = np.concatenate([cols_of_entry, np.array(extra_c, dtype=np.int64)])
    A = sp.coo_matrix((-np.ones(len(I)), (I, J)), shape=(m, n)).tocsc()
    A.sum_duplicates()
    A.data[:] = -1.0
    c = rng.integers(1, 101, size=n).astype(np.float64)
    return MILP(c=c, c0=0.0, A=A, b=-np.ones(m), row_sense=np.full(m, "le"),
                l=np.zeros(n), u=np.ones(n), int_idx=np.arange(n), name=f"setcover_{m}x{n}_s{seed}")


def _power_law_lengths(rng, m: int, nnz: int, exponent: float, max_len: int,
                       n_pinned: int = 8) -> np.ndarray:
    """Row lengths with a Zipf(exponent) shape on {1..max_len}, rescaled so
    that they sum to exactly ``nnz``; the ``n_pinned`` longest rows are
    pinned to ``max_len`` so that the long-row path is exercised."""
    support = np.arange(1, max_len + 1, dtype=np.float64)
    p = support ** (-exponent)
    p /= p.sum()
    K = rng.choice(np.arange(1, max_len + 1), size=m, p=p).astype(np.float64)
    L = np.maximum(1, np.round(K * nnz / K.sum())).astype(np.int64)
    L = np.minimum(L, max_len)
    top = np.argsort(-L, kind="stable")[:n_pinned]
    L[top] = max_len
    # repair the total with ±1 on random non-pinned rows
    free = np.setdiff1d(np.arange(m), top)
    diff = nnz - int(L.sum())
    while diff != 0:
        idx = rng.choice(free, size=min(abs(diff), len(free)), replace=False)
        if diff > 0:
            idx = idx[L[idx] < max_len]
            L[idx] += 1
            diff -= len(idx)
        else:
            idx = idx[L[idx] > 1]
            L[idx] -= 1
            diff += len(idx)
    return L


def uneven_set_covers(count: int = 64, *, m_range=(2_000, 40_000), n_range=(4_000, 80_000),
                      first_seed: int = 1) -> list:
    """``count`` set-covering instances of uneven sizes, seeds first_seed..:
    instance s has m = round(e^U[ln m_lo, ln m_hi]) rows and
    n = round(e^U[ln n_lo, ln n_hi]) columns, drawn with its own seed, and is
    ``set_cover(m, n, s)``."""
    out = []
    for s in range(first_seed, first_seed + count):
        rng = _rng(s)
        m = int(round(float(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1]))))))
        n = int(round(float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))))
        out.append(set_cover(m, n, s))
    return out


def mixed(m: int = 50_000, n: int = 100_000, seed: int = 1, *, min_len: int = 10,
          max_len: int = 30, nnz: int | None = None, power_law: bool = False,
          zipf_exponent: float = 1.5, binary_fraction: float = 0.7,
          cont_ub: float = 10.0) -> MILP:
    """Mixed binary/continuous MILP with N(0,1) values, all rows ≤,
    b = A x0 + |N(0,1)| for a random x0 (binary 0/1, continuous U[0, ub])."""
    rng = _rng(seed)
    if power_law:
        if nnz is None:
            raise ValueError("power_law needs nnz")
        lengths = _power_law_lengths(rng, m, nnz, zipf_exponent, max_len)
    else:
        lengths = rng.integers(min_len, max_len + 1, size=m)
    ptr, cols = _distinct_columns(rng, lengths, n)
    vals = rng.standard_normal(len(cols))
    A = _csr_to_csc(m, n, ptr, cols, vals)
    is_bin = rng.random(n) < binary_fraction
    int_idx = np.flatnonzero(is_bin)
    u = np.where(is_bin, 1.0, cont_ub)
    x0 = np.where(is_bin, rng.integers(0, 2, size=n).astype(np.float64), rng.random(n) * cont_ub)
    b = A @ x0 + np.abs(rng.standard_normal(m))
    return MILP(c=np.zeros(n), c0=0.0, A=A, b=b, row_sense=np.full(m, "le"), l=np.zeros(n),
                u=u, int_idx=int_idx, name=f"mixed_{m}x{n}_s{seed}")


def general_integer(m: int = 200_000, n: int = 400_000, seed: int = 1, *,
                    eq_fraction: float = 0.4, min_len: int = 10, max_len: int = 30,
                    max_ub: int = 20, max_coef: int = 9) -> MILP:
    """General-integer, equality-heavy MILP."""
    rng = _rng(seed)
    lengths = rng.integers(min_len, max_len + 1, size=m)
    ptr, cols = _distinct_columns(rng, lengths, n)
    mag = rng.integers(1, max_coef + 1, size=len(cols))
    sgn = np.where(rng.random(len(cols)) < 0.5, -1, 1)
    vals = (mag * sgn).astype(np.float64)
    A = _csr_to_csc(m, n, ptr, cols, vals)
    U = rng.integers(1, max_ub + 1, size=n).astype(np.float64)
    x0 = np.floor(rng.random(n) * (U + 1.0))
    is_eq = rng.random(m) < eq_fraction
    Ax0 = A @ x0
    b = np.where(is_eq, Ax0, Ax0 + np.abs(rng.standard_normal(m)))
    return MILP(c=np.zeros(n), c0=0.0, A=A, b=b, row_sense=np.where(is_eq, "eq", "le"),
                l=np.zeros(n), u=U, int_idx=np.arange(n), name=f"genint_{m}x{n}_s{seed}")


def binary_targets(milp: MILP, seed: int) -> np.ndarray:
    """A pump-like target x̃: uniform 0/1 on binaries, a uniform integer in
    [l, u] on general integers, and 0 (clamped into the box) elsewhere."""
    rng = _rng(seed)
    n = len(milp.c)
    x = np.clip(np.zeros(n), milp.l, milp.u)
    I = milp.int_idx
    lo = np.where(np.isfinite(milp.l[I]), milp.l[I], -10.0)
    hi = np.where(np.isfinite(milp.u[I]), milp.u[I], lo + 10.0)
    x[I] = np.floor(lo + rng.random(len(I)) * (hi - lo + 1.0))
    x[I] = np.minimum(x[I], hi)
    return x


def toy1() -> MILP:
    """min |x1 - 1| + |x2 - 1| s.t. x1 + x2 ≤ 1, 0 ≤ x ≤ 1 (test/runtests.jl:42-43):
    the ℓ1 projection of (1, 1) has distance 1."""
    return MILP(c=np.zeros(2), c0=0.0, A=sp.csc_matrix(np.array([[1.0, 1.0]])), b=np.array([1.0]),
                row_sense=np.array(["le"]), l=np.zeros(2), u=np.ones(2), int_idx=np.array([0, 1]),
                var_names=["x1", "x2"], name="toy")


def toy2() -> MILP:
    """min |x1 - 3| s.t. x1 + x2 = 2, 0 ≤ x1 ≤ 5, x2 ≥ 0 continuous
    (test/runtests.jl:50-51): the projection of (3, 0) has x1 = 2."""
    return MILP(c=np.zeros(2), c0=0.0, A=sp.csc_matrix(np.array([[1.0, 1.0]])), b=np.array([2.0]),
                row_sense=np.array(["eq"]), l=np.zeros(2), u=np.array([5.0, np.inf]),
                int_idx=np.array([0]), var_names=["x1", "x2"], name="toy2")
