"""CPU pins of the oracle's residuals in original units and of its
infeasibility verdicts, against closed forms evaluated on the unscaled model.

The projector measures its primal residual in original units on mixed models
(``original_primal_residual!``, src/pdfp.jl:85-90, used at src/pdfp.jl:200-205),
and the root LP measures both residuals that way under cuPDLPx termination
(``cupdlpx_termination!``, src/pdlp.jl:105-110, 220-242, 292-316).  With the
scaled problem A_s = D_r A D_c, b_s = D_r b (and, for the root LP, bound scale
β and objective scale γ), every such residual equals a formula on the
original A, b, c and the unscaled point:

- projector:  x = D_c x̂;  ‖viol(A x − b)‖₂ / (1 + ‖b‖₂)
- root LP:    x = D_c x̂ / β,  y = D_r ŷ / γ,  λ = λ̂ / (D_c γ);
  primal ‖A x − Π_[lo,hi](A x)‖₂ / (1 + ‖bounds‖₂),
  dual ‖c − Aᵀ y − λ‖₂ / (1 + ‖c‖₂),
  gap |cᵀx − (λᵀx + Σ max(y,0) lo + min(y,0) hi)| / (1 + |cᵀx| + |dual obj|).

These sums are written here directly on the original data (numpy/scipy), so
an unscaling slip in the oracle — D_r where 1/D_r belongs, ‖b‖ taken from the
scaled b, a column scaling applied to the rows — fails them.  The verdicts
PRIMAL_INFEASIBLE and DUAL_INFEASIBLE (src/pdlp.jl:289-290, 355-362) are
pinned against HiGHS's status on the same LPs.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd.milp import MILP, read_mps
from pdfp_amd.pdfp import SolverParameters, projector_problem

from .conftest import example_path


def _models():
    pk1 = read_mps(example_path("pk1"))
    return {"pk1": pk1, "mixed300x600": G.mixed(300, 600, 2),
            "genint_eq": G.general_integer(200, 400, 3)}


def _viol(milp, x):
    r = milp.A @ x - milp.b
    return np.where(milp.row_sense == "eq", np.abs(r), np.maximum(r, 0.0))


# ---- projector: primal residual in original units (src/pdfp.jl:85-90, 200-205) ----

@pytest.mark.parametrize("name", ["pk1", "mixed300x600", "genint_eq"])
def test_projector_original_units_primal_residual(name):
    milp = _models()[name]
    pp = projector_problem(milp, scaling="host")
    s = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    O.original_primal_residual(s, pp.D_r)
    rng = np.random.default_rng(7)
    n, m = len(milp.c), len(milp.b)
    xt = np.round(rng.uniform(milp.l, np.minimum(milp.u, milp.l + 5.0)))
    O.set_target(s, xt / pp.D_c)
    s.initialize(SolverParameters(), 1.0)
    for trial in range(4):
        # points in original units, well outside and near the feasible set
        x = rng.uniform(milp.l, np.minimum(milp.u, milp.l + 10.0)) * (0.2 + 0.4 * trial)
        xh = x / pp.D_c
        y, z, ds = rng.standard_normal(n), np.abs(rng.standard_normal(m)), rng.standard_normal(n)
        got = s.residual_of(xh, y, z, ds)
        want = np.linalg.norm(_viol(milp, x)) / (1.0 + np.linalg.norm(milp.b))
        assert want > 0
        assert abs(got[0] - want) <= 1e-12 * want, (got[0], want)
        # the ℓ1 distance is Σ_I |x − x̃| in original units (w = D_c on I)
        I = milp.int_idx
        l1 = np.abs(x[I] - xt[I]).sum()
        assert abs(got[3] - l1) <= 1e-12 * l1


@pytest.mark.parametrize("name", ["pk1", "mixed300x600"])
def test_projector_scaled_residuals(name):
    """Without original_primal_residual! the primal residual is on the scaled
    rows, ‖D_r viol(A x − b)‖ / (1 + ‖D_r b‖); the dual residual and the gap
    follow src/pdfp.jl:206-216 with A_s = diag(D_r) A diag(D_c)."""
    milp = _models()[name]
    pp = projector_problem(milp, scaling="host")
    s = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    rng = np.random.default_rng(11)
    n, m = len(milp.c), len(milp.b)
    xt = np.round(rng.uniform(milp.l, np.minimum(milp.u, milp.l + 5.0)))
    O.set_target(s, xt / pp.D_c)
    s.initialize(SolverParameters(), 1.0)
    I = milp.int_idx
    is_int = np.zeros(n, dtype=bool)
    is_int[I] = True
    As = sp.diags(pp.D_r) @ milp.A @ sp.diags(pp.D_c)
    for _ in range(3):
        x = rng.uniform(milp.l, np.minimum(milp.u, milp.l + 10.0))
        y = np.where(is_int, rng.standard_normal(n), 0.0)
        z, ds = np.abs(rng.standard_normal(m)), rng.standard_normal(n)
        got = s.residual_of(x / pp.D_c, y, z, ds)
        prim = np.linalg.norm(pp.D_r * _viol(milp, x)) / (1.0 + np.linalg.norm(pp.D_r * milp.b))
        assert abs(got[0] - prim) <= 1e-12 * prim
        dual = np.linalg.norm(As.T @ z - y - ds) / (1.0 + np.linalg.norm(pp.D_c[I]))
        assert abs(got[1] - dual) <= 1e-11 * dual
        l1 = np.abs(x[I] - xt[I]).sum()
        yterm = float(np.sum(y[I] * (xt[I] - x[I]) / pp.D_c[I]))
        zterm = float(z @ (pp.D_r * (milp.A @ x - milp.b)))
        gap = abs(yterm + zterm - l1) / (1.0 + np.abs(xt[I]).sum())
        assert abs(got[2] - gap) <= 1e-10 * max(gap, 1.0)


# ---- root LP: cuPDLPx residuals in original units (src/pdlp.jl:220-242, 292-316) ----

def _lp_state(milp, cupdlpx):
    con_lb = np.where(milp.row_sense == "le", -np.inf, milp.b)
    spd = O.scale_problem(milp.A, milp.c.copy(), milp.l.copy(), milp.u.copy(), con_lb,
                          milp.b.copy())
    s = O.OraclePDLPState(spd)
    if cupdlpx:
        s.cupdlpx_termination(spd["D_r"], spd["D_c"])
    return s, spd, con_lb


@pytest.mark.parametrize("iters", [200, 600])
@pytest.mark.parametrize("name", ["pk1", "mixed300x600"])
def test_root_lp_cupdlpx_residuals_in_original_units(name, iters):
    milp = _models()[name]
    if name == "mixed300x600":
        milp.c = np.random.default_rng(3).standard_normal(len(milp.c))
    s, spd, con_lb = _lp_state(milp, True)
    p = SolverParameters(eps_optimal_relative=1e-12, eps_feasible_relative=1e-12,
                         iteration_limit=iters)
    assert s.solve(p) == "ITERATION_LIMIT"
    info = s.info()
    bs, os_ = spd["bound_scale"], spd["objective_scale"]
    x = spd["D_c"] * s.vector("pdhg_primal") / bs
    y = spd["D_r"] * s.vector("pdhg_dual") / os_
    lam = s.vector("dual_slack") / (spd["D_c"] * os_)
    lo, hi, c, A = con_lb, milp.b, milp.c, milp.A
    Ax = A @ x
    prim = np.linalg.norm(Ax - np.clip(Ax, lo, hi))
    bnd = np.sqrt(np.sum(np.where(np.isfinite(lo) & (lo != hi), lo, 0.0) ** 2) +
                  np.sum(np.where(np.isfinite(hi), hi, 0.0) ** 2))
    want_p = prim / (1.0 + bnd)
    want_d = np.linalg.norm(c - A.T @ y - lam) / (1.0 + np.linalg.norm(c))
    pobj = float(c @ x)
    dobj = float(lam @ x) + float(np.sum(np.maximum(y, 0.0) * np.where(np.isfinite(lo), lo, 0.0) +
                                         np.minimum(y, 0.0) * np.where(np.isfinite(hi), hi, 0.0)))
    want_g = abs(pobj - dobj) / (1.0 + abs(pobj) + abs(dobj))
    assert want_p > 0 and want_d > 0
    assert abs(info.relative_primal_residual - want_p) <= 1e-9 * want_p
    assert abs(info.relative_dual_residual - want_d) <= 1e-9 * want_d
    assert abs(info.relative_optimality_gap - want_g) <= 1e-9 * max(want_g, 1e-6)
    assert abs(info.objective_value - pobj) <= 1e-9 * (1.0 + abs(pobj))


def test_root_lp_plain_residuals_are_scaled():
    """Without cuPDLPx termination the residuals stay in scaled units
    (src/pdlp.jl:222-233): the primal one against ‖scaled bounds‖ over β."""
    milp = _models()["pk1"]
    s, spd, con_lb = _lp_state(milp, False)
    p = SolverParameters(eps_optimal_relative=1e-12, eps_feasible_relative=1e-12,
                         iteration_limit=200)
    assert s.solve(p) == "ITERATION_LIMIT"
    info = s.info()
    bs = spd["bound_scale"]
    xh = s.vector("pdhg_primal")
    Axh = spd["A"] @ xh
    lo, hi = spd["con_lb"], spd["con_ub"]
    bnd = np.sqrt(np.sum(np.where(np.isfinite(lo) & (lo != hi), lo, 0.0) ** 2) +
                  np.sum(np.where(np.isfinite(hi), hi, 0.0) ** 2))
    want = np.linalg.norm(Axh - np.clip(Axh, lo, hi)) / bs / (1.0 + bnd)
    assert abs(info.relative_primal_residual - want) <= 1e-10 * want


# ---- infeasibility verdicts (src/pdlp.jl:247-290, 355-362) ---------------------

def _highs_status(milp):
    """HiGHS on the LP relaxation: 0 optimal, 2 infeasible, 3 unbounded."""
    A = milp.A.toarray()
    le = milp.row_sense == "le"
    res = linprog(milp.c, A_ub=A[le] if le.any() else None, b_ub=milp.b[le] if le.any() else None,
                  A_eq=A[~le] if (~le).any() else None, b_eq=milp.b[~le] if (~le).any() else None,
                  bounds=list(zip(milp.l, [None if not np.isfinite(u) else u for u in milp.u])),
                  method="highs")
    return res.status


def _unbounded_lp():
    # min -x1 - x2 s.t. x1 - x2 ≤ 1, -x1 + x2 ≤ 1, x ≥ 0: the ray (1, 1)
    A = sp.csc_matrix(np.array([[1.0, -1.0], [-1.0, 1.0]]))
    return MILP(c=np.array([-1.0, -1.0]), c0=0.0, A=A, b=np.array([1.0, 1.0]),
                row_sense=np.array(["le", "le"]), l=np.zeros(2), u=np.array([np.inf, np.inf]),
                int_idx=np.array([0]))


def _unbounded_mixed():
    # a generated mixed model with one continuous column that improves the
    # objective without bound and appears in no row
    m = G.mixed(200, 400, 3)
    cont = np.setdiff1d(np.arange(400), m.int_idx)
    j = int(cont[0])
    A = m.A.tolil()
    A[:, j] = 0.0
    m.A = A.tocsc()
    m.A.eliminate_zeros()
    m.c = np.zeros(400)
    m.u[j] = np.inf
    m.c[j] = -1.0
    return m


def _infeasible_lp():
    A = sp.csc_matrix(np.array([[1.0, 1.0], [-1.0, -1.0]]))
    return MILP(c=np.array([1.0, 1.0]), c0=0.0, A=A, b=np.array([1.0, -3.0]),
                row_sense=np.array(["le", "le"]), l=np.zeros(2), u=np.array([10.0, np.inf]),
                int_idx=np.array([0]))


@pytest.mark.parametrize("maker,status,highs", [(_unbounded_lp, "DUAL_INFEASIBLE", 3),
                                                (_unbounded_mixed, "DUAL_INFEASIBLE", 3),
                                                (_infeasible_lp, "PRIMAL_INFEASIBLE", 2)])
def test_root_lp_infeasibility_verdicts(maker, status, highs):
    milp = maker()
    assert _highs_status(milp) == highs
    x, ds, st, it = O.solve_lp_relaxation(milp, tolerance=1e-4, cupdlpx_termination=True)
    assert st == status
    # bounding the improving direction makes the same LP optimal
    if status == "DUAL_INFEASIBLE":
        milp.u = np.where(np.isfinite(milp.u), milp.u, 50.0)
        assert _highs_status(milp) == 0
        x, ds, st, it = O.solve_lp_relaxation(milp, tolerance=1e-4, cupdlpx_termination=True)
        assert st == "OPTIMAL"
