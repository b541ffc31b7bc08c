"""Random tiny MILPs for the property tests (CPU and GPU).

Each instance is feasible by construction: b = A x0 + slack (≤ rows) or
b = A x0 (= rows) for a random x0 inside the bounds, with integer columns
binary or general (upper bounds 1-3), continuous columns in [0, U], integer
coefficients in [-5, 5], equality rows, and rows or columns left empty when
the draw leaves them so."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from hypothesis import strategies as st

from pdfp_amd.milp import MILP


def tiny_milp(m: int, n: int, density: float, eq_frac: float, int_frac: float,
              seed: int) -> MILP:
    rng = np.random.default_rng(seed)
    mask = rng.random((m, n)) < density
    vals = rng.integers(-5, 6, size=(m, n)).astype(np.float64)
    vals[vals == 0] = 1.0
    A = sp.csc_matrix(np.where(mask, vals, 0.0))
    is_int = rng.random(n) < int_frac
    u = np.where(is_int, rng.integers(1, 4, size=n).astype(np.float64),
                 rng.uniform(1.0, 10.0, size=n))
    x0 = np.where(is_int, np.floor(rng.random(n) * (u + 1)), rng.random(n) * u)
    eq = rng.random(m) < eq_frac
    b = A @ x0 + np.where(eq, 0.0, np.abs(rng.standard_normal(m)))
    return MILP(c=np.zeros(n), c0=0.0, A=A, b=b, row_sense=np.where(eq, "eq", "le"),
                l=np.zeros(n), u=u, int_idx=np.flatnonzero(is_int),
                name=f"tiny_{m}x{n}_s{seed}")


def integer_target(milp: MILP, seed: int) -> np.ndarray:
    """x̃: random integers inside the bounds on the integer columns, 0 elsewhere."""
    rng = np.random.default_rng(seed + 7919)
    xt = np.zeros(len(milp.c))
    I = milp.int_idx
    xt[I] = np.floor(rng.random(len(I)) * (milp.u[I] + 1))
    return xt


def fractional_point(milp: MILP, seed: int) -> np.ndarray:
    """A point inside the bounds, fractional on about half of the integer
    columns (the others integral), as an LP point handed to the rounder."""
    rng = np.random.default_rng(seed + 104729)
    x = rng.random(len(milp.c)) * milp.u
    I = milp.int_idx
    keep = rng.random(len(I)) < 0.5
    x[I[keep]] = np.rint(x[I[keep]])
    return x


# (m, n, density, eq_frac, int_frac, seed)
instances = st.tuples(st.integers(2, 12), st.integers(3, 20), st.floats(0.15, 0.7),
                      st.sampled_from([0.0, 0.25, 0.5]), st.sampled_from([0.4, 0.7, 1.0]),
                      st.integers(0, 10_000))
