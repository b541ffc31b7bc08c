"""CPU tests that pin the oracle (the FP64 C restatement of the reference's
CPU path) and the host-side pieces that stay on the CPU.

Pins, per oracle function:
- projector solve (src/pdfp.jl:252-288): the reference's closed-form toys
  (test/runtests.jl:40-55) and HiGHS ℓ1 optima of 15 instances
  (tests/golden/l1_projections.json, made by tests/golden/make_fixtures.py).
- scale_problem (src/preconditioner.jl:26-99): bit-identical to the numpy
  restatement; A_s = diag(D_r) A diag(D_c); closed forms on a scaled
  permutation matrix and on {0, ±1} matrices (Ruiz is the identity there and
  Pock–Chambolle divides by the square roots of the row and column counts).
- estimate_max_singular_value (src/preconditioner.jl:108-125): dense SVD.
- CSC product (SparseArrays mul!): scipy's CSC product.
- pdlp_solve (src/pdlp.jl:325-372): HiGHS LP optima (tests/golden/root_lp.json),
  p0201's published LP value 6875 (examples/p0201.mps.gz header) and the
  README's 600 root-LP iterations on p0201 (README.md:39), exactly.
"""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd.milp import is_feasible, max_violation, num_continuous, read_mps
from pdfp_amd.pdfp import SolverParameters
from pdfp_amd.preconditioner import scale_problem

from .conftest import example_path
from .helpers import build_instance, l1_distance, l1_fixtures, project_with, root_lp_fixtures


# ---- MPS reader (test/runtests.jl:30-38) -------------------------------------

def test_mps_reader_p0201_pk1():
    m = read_mps(example_path("p0201"))
    assert len(m.c) == 201
    assert len(m.int_idx) == 201
    assert np.all(m.l[m.int_idx] == 0) and np.all(m.u[m.int_idx] == 1)
    assert set(m.row_sense) <= {"le", "eq"}
    assert m.A.shape == (133, 201) and m.A.nnz == 1923   # README.md:36-37
    m = read_mps(example_path("pk1"))
    assert num_continuous(m) == 31


def test_max_violation_and_is_feasible():
    m = G.toy2()   # x1 + x2 = 2, x1 in [0,5] integer, x2 >= 0
    assert max_violation(m, [2.0, 0.0]) == 0.0
    assert max_violation(m, [1.0, 0.0]) == 1.0
    assert is_feasible(m, [2.0, 0.0])
    assert not is_feasible(m, [1.5, 0.5])         # integrality
    assert not is_feasible(m, [6.0, -4.0])        # bounds
    t = G.toy1()
    assert max_violation(t, [1.0, 1.0]) == 1.0
    assert max_violation(t, [0.0, 0.0]) == 0.0    # slack on a ≤ row is no violation


# ---- projector: known answers --------------------------------------------------

def test_projection_toys():
    """test/runtests.jl:40-55 on the oracle."""
    pp, s, status = project_with(O, O.OraclePDFPState, G.toy1(), [1.0, 1.0])
    x = pp.D_c * s.pdhg_x
    assert status == "OPTIMAL"
    assert abs(np.abs(x - 1.0).sum() - 1.0) <= 1e-5
    assert x.sum() <= 1.0 + 1e-6
    pp, s, status = project_with(O, O.OraclePDFPState, G.toy2(), [3.0, 0.0])
    x = pp.D_c * s.pdhg_x
    assert status == "OPTIMAL"
    assert abs(x[0] - 2.0) <= 1e-5


@pytest.mark.parametrize("case", l1_fixtures(), ids=lambda c: c["name"])
def test_oracle_l1_matches_highs(case):
    milp = build_instance(case["spec"])
    xt = np.array(case["target"])
    pp, s, status = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-8)
    assert status == "OPTIMAL"
    x = pp.D_c * s.pdhg_x
    ref = case["l1_optimum"]
    assert abs(l1_distance(milp, x, xt) - ref) <= 1e-6 * max(1.0, ref)
    assert max_violation(milp, x) <= 1e-5 * (1 + np.abs(milp.b).max())


def test_single_row_fractional_knapsack():
    """min Σ|x_j - 1| s.t. Σ a_j x_j ≤ β, 0 ≤ x ≤ 1: move the cheapest
    coordinates (largest a_j per unit of ℓ1) down first — a closed form."""
    a = np.array([5.0, 1.0, 3.0, 2.0])
    beta = 4.0
    A = sp.csc_matrix(a[None, :])
    milp = G.MILP(c=np.zeros(4), c0=0.0, A=A, b=np.array([beta]), row_sense=np.array(["le"]),
                  l=np.zeros(4), u=np.ones(4), int_idx=np.arange(4))
    # need to remove Σa - β = 7 units of activity; largest a first: x1 → 0 (5), x3 → 1/3 (2)
    expected = 1.0 + 2.0 / 3.0
    pp, s, status = project_with(O, O.OraclePDFPState, milp, np.ones(4), tol=1e-9)
    x = pp.D_c * s.pdhg_x
    assert status == "OPTIMAL"
    assert abs(l1_distance(milp, x, np.ones(4)) - expected) <= 1e-6


def test_projection_of_feasible_target_is_itself():
    m = G.set_cover(133, 201, 5)
    xt = np.ones(201)                 # covers every row
    pp, s, status = project_with(O, O.OraclePDFPState, m, xt, tol=1e-8)
    x = pp.D_c * s.pdhg_x
    assert status == "OPTIMAL"
    assert l1_distance(m, x, xt) <= 1e-6


def test_no_rows_projection_is_box_clamp():
    """With no constraints, the projection is clamp(x̃, l, u)."""
    n = 6
    A = sp.csc_matrix((0, n))
    xt = np.array([-1.0, 0.5, 2.0, 3.0, 0.0, 1.0])
    milp = G.MILP(c=np.zeros(n), c0=0.0, A=A, b=np.zeros(0), row_sense=np.zeros(0, dtype="<U2"),
                  l=np.zeros(n), u=np.full(n, 1.0), int_idx=np.arange(n))
    pp, s, status = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-9)
    x = pp.D_c * s.pdhg_x
    assert status == "OPTIMAL"
    np.testing.assert_allclose(x, np.clip(xt, 0.0, 1.0), atol=1e-7)


def test_iteration_and_time_limits():
    m = G.set_cover(133, 201, 1)
    pp, s, status = project_with(O, O.OraclePDFPState, m, np.zeros(201), tol=1e-14,
                                 iteration_limit=400)
    assert status == "ITERATION_LIMIT" and s.total_count == 400
    O.warm_start(s)
    p = SolverParameters(eps_optimal_relative=1e-14, eps_feasible_relative=1e-14,
                         time_limit=0.0)
    assert O.pdfp_solve(s, p, 1.0) == "TIME_LIMIT"
    assert s.total_count == 200


# ---- scaling ---------------------------------------------------------------------

@pytest.mark.parametrize("rescale", [False, True])
@pytest.mark.parametrize("maker", [lambda: G.set_cover(133, 201, 2),
                                   lambda: G.mixed(500, 900, 4),
                                   lambda: read_mps(example_path("pk1")),
                                   lambda: read_mps(example_path("gen-ip054"))])
def test_scaling_numpy_equals_oracle_bitwise(maker, rescale):
    m = maker()
    con_lb = np.where(m.row_sense == "eq", m.b, -np.inf)
    a = scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, bound_objective_rescaling=rescale,
                      backend="host")
    b = O.scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, bound_objective_rescaling=rescale)
    for k in ("D_r", "D_c", "c", "var_lb", "var_ub", "con_lb", "con_ub"):
        assert np.array_equal(getattr(a, k), b[k]), k
    assert np.array_equal(a.A.data, b["A"].data)
    assert a.bound_scale == b["bound_scale"] and a.objective_scale == b["objective_scale"]
    # invariant: A_s = diag(D_r) A diag(D_c)
    ref = sp.diags(a.D_r) @ m.A @ sp.diags(a.D_c)
    np.testing.assert_allclose(a.A.toarray(), ref.toarray(), rtol=1e-13, atol=0)


def test_scaling_permutation_closed_form():
    """One nonzero a_k per row and column: Ruiz makes every entry ±1 in the
    first round (a / (√|a|·√|a|)), Pock–Chambolle then leaves it, so
    D_r[i]·D_c[j] = 1/|a|."""
    vals = np.array([4.0, -0.25, 9.0, 2.0])
    perm = np.array([2, 0, 3, 1])
    A = sp.csc_matrix((vals, (np.arange(4), perm)), shape=(4, 4))
    a = scale_problem(A, np.zeros(4), np.zeros(4), np.ones(4), np.full(4, -np.inf), np.ones(4),
                      bound_objective_rescaling=False, backend="host")
    np.testing.assert_allclose(np.abs(a.A.data), 1.0, rtol=1e-15)
    np.testing.assert_allclose(a.D_r * a.D_c[perm], 1.0 / np.abs(vals), rtol=1e-15)


@pytest.mark.parametrize("signs", [False, True])
def test_scaling_unit_entries_closed_form(signs):
    """Entries in {0, ±1}: every row and column ∞-norm is 1, so Ruiz leaves
    the matrix alone, and Pock–Chambolle with α = 1 divides row i by
    √(nnz in row i) and column j by √(nnz in column j); rows and columns
    without nonzeros keep scale 1 (SCALING_EPSILON, src/preconditioner.jl:38-60)."""
    m = G.set_cover(133, 201, 4)
    A = m.A.tocsc().astype(np.float64)
    if signs:
        A.data = np.where(np.random.default_rng(1).random(A.nnz) < 0.5, -1.0, 1.0)
    # one empty row and one empty column
    A = sp.vstack([A, sp.csr_matrix((1, A.shape[1]))]).tocsc()
    A = sp.hstack([A, sp.csc_matrix((A.shape[0], 1))]).tocsc()
    mm, nn = A.shape
    b = O.scale_problem(A, np.zeros(nn), np.zeros(nn), np.ones(nn), np.full(mm, -np.inf),
                        np.ones(mm), bound_objective_rescaling=False)
    rc = np.diff(A.tocsr().indptr).astype(float)
    cc = np.diff(A.indptr).astype(float)
    D_r = np.where(rc > 0, 1.0 / np.sqrt(np.maximum(rc, 1.0)), 1.0)
    D_c = np.where(cc > 0, 1.0 / np.sqrt(np.maximum(cc, 1.0)), 1.0)
    np.testing.assert_allclose(b["D_r"], D_r, rtol=1e-15)
    np.testing.assert_allclose(b["D_c"], D_c, rtol=1e-15)
    assert b["D_r"][-1] == 1.0 and b["D_c"][-1] == 1.0
    ref = (sp.diags(D_r) @ A @ sp.diags(D_c)).tocsc()
    ref.sort_indices()
    np.testing.assert_allclose(b["A"].toarray(), ref.toarray(), rtol=1e-15, atol=0)


# ---- power iteration, SpMV ------------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_singular_value_matches_svd(seed):
    rng = np.random.default_rng(seed)
    A = sp.random(40, 70, density=0.2, random_state=seed, format="csc")
    A.data = rng.standard_normal(A.nnz)
    sv = O.max_singular_value(A, 5000, 1e-10)
    ref = np.linalg.svd(A.toarray(), compute_uv=False)[0]
    assert abs(sv - ref) <= 1e-8 * ref


def test_oracle_spmv_matches_scipy():
    m = G.mixed(400, 700, 9)
    x = np.random.default_rng(0).standard_normal(700)
    z = np.random.default_rng(1).standard_normal(400)
    np.testing.assert_allclose(O.spmv(m.A, x), m.A @ x, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(O.spmv(m.A, z, transpose=True), m.A.T @ z, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("unsorted", [False, True])
def test_threaded_products_are_bit_identical(unsorted):
    """The oracle's optional threaded step (orc_set_threads, the all-core
    CPU baseline) sums each row in the scatter's order: a projection
    with 4 threads gives the single-threaded iterates, counts and distance
    bit for bit.  With the caller's row indices out of order within a
    column the threaded Aᵀ z would sum its rows in another order, so that
    product falls back to the scatter (the same bits again)."""
    from pdfp_amd.pdfp import projector_problem

    m = G.mixed(300, 700, 4, power_law=True, nnz=6000, max_len=200)
    pp = projector_problem(m, scaling="host")
    A = pp.A.tocsc()
    if unsorted:  # reverse every column's entries: the same matrix, unsorted
        A = A.copy()
        for j in range(A.shape[1]):
            a, b = A.indptr[j], A.indptr[j + 1]
            A.indices[a:b] = A.indices[a:b][::-1].copy()
            A.data[a:b] = A.data[a:b][::-1].copy()
        A.has_sorted_indices = False
    xt = G.binary_targets(m, 3) / pp.D_c
    p = SolverParameters(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6,
                         iteration_limit=3000)
    runs = []
    for threads in (1, 4):
        O.set_threads(threads)
        try:
            o = O.OraclePDFPState(A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
            O.set_target(o, xt)
            O.pdfp_solve(o, p, 1.0)
            runs.append((o.total_count, o.info().l1_distance, o.pdhg_x, o.vector("pdhg_z")))
        finally:
            O.set_threads(1)
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][2], runs[1][2]) and np.array_equal(runs[0][3], runs[1][3])


def test_start_vector_is_standard_normal():
    v = O.start_vector(200_000)
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1.0) < 0.01
    assert np.array_equal(v, O.start_vector(200_000))


# ---- root LP ------------------------------------------------------------------------------

@pytest.mark.parametrize("case", root_lp_fixtures(), ids=lambda c: c["name"])
def test_oracle_root_lp_matches_highs(case):
    milp = build_instance(case["spec"])
    mixed = num_continuous(milp) > 0
    x, ds, status, iters = O.solve_lp_relaxation(milp, tolerance=1e-6,
                                                 cupdlpx_termination=mixed)
    assert status == "OPTIMAL"
    obj = float(milp.c @ x) + milp.c0
    ref = case["lp_optimum"]
    assert abs(obj - ref) <= 1e-3 * (1.0 + abs(ref))


def test_p0201_root_lp_iterations_exact():
    """README.md:38-39: the root LP of p0201 at tolerance 1e-2 ends OPTIMAL
    after 600 PDHG iterations, which the oracle reproduces exactly."""
    milp = read_mps(example_path("p0201"))
    x, ds, status, iters = O.solve_lp_relaxation(milp, tolerance=1e-2)
    assert status == "OPTIMAL"
    assert iters == 600
    assert abs(float(milp.c @ x) - 6875.0) <= 0.05 * 6875.0


# ---- the product never routes through the oracle ----------------------------------------

def test_product_does_not_import_oracle():
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pdfp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f), encoding="utf-8").read()
                assert not re.search(r"^\s*(from|import)\s+oracle", src, re.M), f
                assert "liboracle" not in src, f
