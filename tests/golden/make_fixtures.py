"""Writes tests/golden/l1_projections.json and tests/golden/root_lp.json.

Known answers that pin the projector and the root LP without the reference:
HiGHS (scipy.optimize.linprog, method="highs") solves, as a plain LP,

    ℓ1 projection:  min Σ_{j∈I} t_j   s.t.  t ≥ x - x̃,  t ≥ x̃ - x,
                                          A x (≤, =) b,  l ≤ x ≤ u
    root LP:        min c'x           s.t.  A x (≤, =) b,  l ≤ x ≤ u

on small generated instances and on the four example instances.  The
projector's weighted ℓ1 (weights D_c in scaled units) equals Σ_I |x - x̃| in
original units, so the optimal values are directly comparable.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pdfp_amd import generators as G  # noqa: E402
from pdfp_amd.milp import read_mps  # noqa: E402


def l1_projection_value(milp, xt) -> float:
    n = len(milp.c)
    I = milp.int_idx
    k = len(I)
    # variables [x (n), t (k)]
    c = np.concatenate([np.zeros(n), np.ones(k)])
    E = sp.csr_matrix((np.ones(k), (np.arange(k), I)), shape=(k, n))
    Ik = sp.identity(k, format="csr")
    # x - t ≤ x̃ ;  -x - t ≤ -x̃
    A_abs = sp.vstack([sp.hstack([E, -Ik]), sp.hstack([-E, -Ik])]).tocsr()
    b_abs = np.concatenate([xt[I], -xt[I]])
    le = milp.row_sense == "le"
    eq = ~le
    A = sp.csr_matrix(milp.A)
    Z = sp.csr_matrix((A.shape[0], k))
    A_le = sp.vstack([sp.hstack([A[le], Z[le]]), A_abs]).tocsr()
    b_le = np.concatenate([milp.b[le], b_abs])
    kw = {}
    if eq.any():
        kw = dict(A_eq=sp.hstack([A[eq], Z[eq]]).tocsr(), b_eq=milp.b[eq])
    bounds = [(lo if np.isfinite(lo) else None, hi if np.isfinite(hi) else None)
              for lo, hi in zip(milp.l, milp.u)] + [(0, None)] * k
    res = linprog(c, A_ub=A_le, b_ub=b_le, bounds=bounds, method="highs", **kw)
    if res.status != 0:
        raise RuntimeError(f"HiGHS: {res.message}")
    return float(res.fun)


def lp_value(milp) -> float:
    le = milp.row_sense == "le"
    eq = ~le
    A = sp.csr_matrix(milp.A)
    kw = {}
    if eq.any():
        kw = dict(A_eq=A[eq], b_eq=milp.b[eq])
    bounds = [(lo if np.isfinite(lo) else None, hi if np.isfinite(hi) else None)
              for lo, hi in zip(milp.l, milp.u)]
    res = linprog(milp.c, A_ub=A[le] if le.any() else None, b_ub=milp.b[le] if le.any() else None,
                  bounds=bounds, method="highs", **kw)
    if res.status != 0:
        raise RuntimeError(f"HiGHS: {res.message}")
    return float(res.fun) + milp.c0


def cases():
    """(name, builder spec, milp, target)"""
    out = []
    out.append(("toy1", {"family": "toy1"}, G.toy1(), np.array([1.0, 1.0])))
    out.append(("toy2", {"family": "toy2"}, G.toy2(), np.array([3.0, 0.0])))
    for seed in (1, 2, 3):
        m = G.set_cover(133, 201, seed)
        xt = np.zeros(201)                       # the empty cover: distance = fractional cover size
        out.append((f"setcover_zero_s{seed}", {"family": "set_cover", "m": 133, "n": 201,
                                               "seed": seed, "target": "zeros"}, m, xt))
        rng = np.random.default_rng(100 + seed)
        xt = (rng.random(201) < 0.15).astype(float)
        out.append((f"setcover_sparse_s{seed}", {"family": "set_cover", "m": 133, "n": 201,
                                                 "seed": seed, "target": "bernoulli0.15",
                                                 "target_seed": 100 + seed}, m, xt))
    for seed in (1, 2):
        m = G.mixed(300, 600, seed)
        xt = G.binary_targets(m, 50 + seed)
        out.append((f"mixed_300x600_s{seed}", {"family": "mixed", "m": 300, "n": 600,
                                               "seed": seed, "target_seed": 50 + seed}, m, xt))
    m = G.general_integer(120, 240, 7)
    xt = G.binary_targets(m, 57)
    out.append(("genint_120x240_s7", {"family": "general_integer", "m": 120, "n": 240, "seed": 7,
                                      "target_seed": 57}, m, xt))
    ex = os.path.join(HERE, "examples")
    for name in ("p0201", "pk1", "gen-ip054", "markshare2"):
        milp = read_mps(os.path.join(ex, name + ".mps.gz"))
        rng = np.random.default_rng(4)
        xt = np.zeros(len(milp.c))
        I = milp.int_idx
        lo = milp.l[I]
        hi = np.where(np.isfinite(milp.u[I]), milp.u[I], lo + 3.0)
        xt[I] = np.round(lo + rng.random(len(I)) * (hi - lo))
        out.append((f"example_{name}", {"family": "example", "name": name, "target_seed": 4},
                    milp, xt))
    return out


def main():
    proj = []
    for name, spec, milp, xt in cases():
        v = l1_projection_value(milp, xt)
        proj.append({"name": name, "spec": spec, "target": xt.tolist(), "l1_optimum": v})
        print(f"{name:28s} ℓ1* = {v:.10g}")
    with open(os.path.join(HERE, "l1_projections.json"), "w") as f:
        json.dump(proj, f, indent=1)
    lps = []
    for name in ("p0201", "pk1", "gen-ip054", "markshare2"):
        milp = read_mps(os.path.join(HERE, "examples", name + ".mps.gz"))
        v = lp_value(milp)
        lps.append({"name": name, "spec": {"family": "example", "name": name}, "lp_optimum": v})
        print(f"LP {name:25s} = {v:.10g}")
    for seed in (1, 2, 3):
        milp = G.set_cover(133, 201, seed)
        v = lp_value(milp)
        lps.append({"name": f"setcover_s{seed}",
                    "spec": {"family": "set_cover", "m": 133, "n": 201, "seed": seed},
                    "lp_optimum": v})
        print(f"LP setcover_s{seed:<17d} = {v:.10g}")
    with open(os.path.join(HERE, "root_lp.json"), "w") as f:
        json.dump(lps, f, indent=1)


if __name__ == "__main__":
    main()
