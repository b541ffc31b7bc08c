"""GPU parity of the projector (libpdfp_hip through its C ABI) against the
oracle and the HiGHS fixtures.

Bars (from the target behaviour of the path):
- FP64 form: the same restarted Halpern PDHG as the oracle, with FP64
  products and sums in a different order only inside long rows and the
  reductions: same termination, same iteration count, iterates within 1e-9
  (scaled units), ℓ1 distance within 1e-9 relative.
- FP32 form: the returned (polished) point's KKT residuals, re-evaluated by
  the oracle's FP64 compute_residual!, meet the requested tolerance to
  within a factor 10 (FP32 rounding of the iterates), and the ℓ1 distance
  matches HiGHS to 1e-4 relative at tolerance 1e-5.  (1e-5 relative KKT is
  about 100 FP32 ulps; tighter tolerances sit at the FP32 noise floor of
  the iterates, which is why the product reaches them in the FP64 form.)
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd import pdfp as P
from pdfp_amd._lib import LibraryError
from pdfp_amd.milp import max_violation, read_mps
from pdfp_amd.pdfp import SolverParameters, projector_problem

from .conftest import example_path
from .helpers import build_instance, l1_distance, l1_fixtures, project_with

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip_state(pp, precision):
    return P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                       precision=precision)


def oracle_state(pp):
    return O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spmv_parity(precision):
    say("spmv parity", precision)
    m = G.mixed(3000, 5000, 11, power_law=True, nnz=60_000, max_len=2500)
    pp = projector_problem(m)
    s = hip_state(pp, precision)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(5000)
    z = rng.standard_normal(3000)
    A = pp.A
    if precision == "fp32":
        A = A.copy()
        A.data = A.data.astype(np.float32).astype(np.float64)
        x = x.astype(np.float32).astype(np.float64)
        z = z.astype(np.float32).astype(np.float64)
    ref = O.spmv(A, x)
    reft = O.spmv(A, z, transpose=True)
    scale = np.abs(A) @ np.abs(x)
    scalet = np.abs(A).T @ np.abs(z)
    got, gott = s.spmv(x), s.spmv(z, transpose=True)
    # products of FP32 operands are exact in FP64 and sums accumulate in FP64
    assert np.all(np.abs(got - ref) <= 1e-14 * (scale + 1e-300))
    assert np.all(np.abs(gott - reft) <= 1e-14 * (scalet + 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_singular_value_parity(precision):
    say("sigma parity", precision)
    m = G.mixed(2000, 3000, 5)
    pp = projector_problem(m)
    s = hip_state(pp, precision)
    sv = s.max_singular_value(5000, 1e-4)
    ref = O.max_singular_value(pp.A, 5000, 1e-4)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert abs(sv - ref) <= tol * ref


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_projection_toys(precision):
    """test/runtests.jl:40-55 through the HIP path."""
    say("toys", precision)
    x, status = P.project(G.toy1(), [1.0, 1.0], tol=1e-8 if precision == "fp64" else 1e-6,
                          precision=precision)
    assert status == "OPTIMAL"
    assert abs(np.abs(x - 1.0).sum() - 1.0) <= 1e-5
    assert x.sum() <= 1.0 + 1e-6
    x, status = P.project(G.toy2(), [3.0, 0.0], tol=1e-8 if precision == "fp64" else 1e-6,
                          precision=precision)
    assert status == "OPTIMAL"
    assert abs(x[0] - 2.0) <= 1e-5


SMALL = [("setcover_s1", lambda: G.set_cover(133, 201, 1), "zeros"),
         ("setcover_s2", lambda: G.set_cover(133, 201, 2), "zeros"),
         ("mixed_300x600", lambda: G.mixed(300, 600, 1), "binary"),
         ("genint_120x240", lambda: G.general_integer(120, 240, 7), "binary"),
         ("powerlaw_2000x4000", lambda: G.mixed(2000, 4000, 3, power_law=True, nnz=40_000,
                                                max_len=3000), "binary")]


def _target(m, kind):
    return np.zeros(len(m.c)) if kind == "zeros" else G.binary_targets(m, 77)


def fp64_state(path):
    """A PDFPState constructor that sets the kernel path before any solve."""
    def make(*a, **kw):
        s = P.PDFPState(*a, **kw)
        s.set_kernel_path(path)
        return s
    return make


def fp64_fits_one_workgroup(pp) -> bool:
    # MI355X: 160 KiB of LDS per workgroup, less 8 KiB for static LDS
    m_, n_ = pp.A.shape
    return (m_ > 0 and n_ > 0 and pp.A.nnz > 0 and
            4 * (n_ + 1) + 4 * (m_ + 1) + 24 * pp.A.nnz + 72 * n_ + 25 * m_ <= 152 * 1024)


@pytest.mark.parametrize("path", ["auto", "one-workgroup"])
@pytest.mark.parametrize("name,maker,kind", SMALL, ids=[c[0] for c in SMALL])
def test_fp64_trajectory_matches_oracle(name, maker, kind, path):
    """The FP64 form on the reference's launch structure ("auto") and with
    whole blocks and the power iteration in one workgroup ("one-workgroup",
    on the instances that fit in LDS): the same bar."""
    say("fp64 trajectory", name, path)
    m = maker()
    xt = _target(m, kind)
    pp, s, st = project_with(P, fp64_state(path), m, xt, tol=1e-7, precision="fp64")
    pq, o, so = project_with(O, O.OraclePDFPState, m, xt, tol=1e-7)
    assert s.info().small_path == (1 if path == "one-workgroup" and fp64_fits_one_workgroup(pp)
                                   else 0)
    assert st == so == "OPTIMAL"
    assert s.total_count == o.total_count
    np.testing.assert_allclose(s.pdhg_x, o.pdhg_x, rtol=0, atol=1e-9)
    np.testing.assert_allclose(s.vector("pdhg_z"), o.vector("pdhg_z"), rtol=0, atol=1e-9)
    li, lo = s.info().l1_distance, o.info().l1_distance
    assert abs(li - lo) <= 1e-9 * max(1.0, lo)
    # the device's FP64 reductions agree with the oracle's sequential sums
    i, j = s.info(), o.info()
    for f in ("relative_primal_residual", "relative_dual_residual", "relative_l1_gap"):
        a, b = getattr(i, f), getattr(j, f)
        assert abs(a - b) <= 1e-6 * abs(b) + 1e-14, f


@pytest.mark.parametrize("path", ["auto", "one-workgroup"])
@pytest.mark.parametrize("case", l1_fixtures(), ids=lambda c: c["name"])
def test_fp64_l1_matches_highs(case, path):
    say("fp64 highs", case["name"], path)
    milp = build_instance(case["spec"])
    xt = np.array(case["target"])
    pp, s, status = project_with(P, fp64_state(path), milp, xt, tol=1e-8, precision="fp64")
    assert s.info().small_path == (1 if path == "one-workgroup" and fp64_fits_one_workgroup(pp)
                                   else 0)
    assert status == "OPTIMAL"
    x = pp.D_c * s.pdhg_x
    ref = case["l1_optimum"]
    assert abs(l1_distance(milp, x, xt) - ref) <= 1e-6 * max(1.0, ref)


@pytest.mark.parametrize("path", ["auto", "general"])
@pytest.mark.parametrize("case", l1_fixtures(), ids=lambda c: c["name"])
def test_fp32_l1_matches_highs_and_kkt(case, path):
    """Both kernel paths of the FP32 form: "auto" runs whole blocks in one
    workgroup wherever the operators fit in LDS."""
    say("fp32 highs", case["name"], path)
    milp = build_instance(case["spec"])
    xt = np.array(case["target"])
    tol = 1e-5
    pp = projector_problem(milp)
    s = hip_state(pp, "fp32")
    s.set_kernel_path(path)
    P.set_target(s, xt / pp.D_c)
    status = P.pdfp_solve(s, SolverParameters(eps_optimal_relative=tol,
                                              eps_feasible_relative=tol,
                                              iteration_limit=200_000), 1.0)
    assert status == "OPTIMAL"
    m_, n_ = pp.A.shape
    # MI355X: 160 KiB of LDS per workgroup, less 8 KiB for static LDS
    fits = 4 * (n_ + 1) + 4 * (m_ + 1) + 16 * pp.A.nnz + 36 * n_ + 13 * m_ <= 152 * 1024
    assert s.info().small_path == (1 if path == "auto" and fits and pp.A.nnz > 0 else 0)
    x = pp.D_c * s.pdhg_x
    ref = case["l1_optimum"]
    assert abs(l1_distance(milp, x, xt) - ref) <= 1e-4 * max(1.0, ref)
    # FP64 KKT of the returned point, evaluated by the oracle
    o = oracle_state(pp)
    O.set_target(o, xt / pp.D_c)
    o.initialize(None, 1.0)
    r = o.residual_of(s.pdhg_x, s.vector("pdhg_y"), s.vector("pdhg_z"), s.vector("dual_slack"))
    assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, r


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_warm_started_sequence_like_the_pump(precision):
    """Three successive warm-started projections with the pump's tolerance
    schedule (src/pump.jl:329-358) — device vs oracle."""
    say("warm sequence", precision)
    m = G.mixed(400, 800, 21)
    pp = projector_problem(m)
    s = hip_state(pp, precision)
    o = oracle_state(pp)
    P.original_primal_residual(s, pp.D_r)
    O.original_primal_residual(o, pp.D_r)
    x_start = np.clip(np.full(800, 0.5), m.l, m.u) / pp.D_c
    s.set_x(x_start)
    o.set_x(x_start)
    gap = 1e-2
    pw = np.sqrt(1.0 + len(m.int_idx) / 400)
    for k in range(3):
        xt = G.binary_targets(m, 300 + k) / pp.D_c
        P.set_target(s, xt)
        O.set_target(o, xt)
        p = SolverParameters(eps_optimal_relative=gap, eps_feasible_relative=gap * 0.1,
                             iteration_limit=500_000)
        ts, to = P.pdfp_solve(s, p, pw), O.pdfp_solve(o, p, pw)
        assert ts == to == "OPTIMAL"
        if precision == "fp64":
            assert s.total_count == o.total_count
            np.testing.assert_allclose(s.pdhg_x, o.pdhg_x, atol=1e-8)
        else:
            r = o.residual_of(s.pdhg_x, s.vector("pdhg_y"), s.vector("pdhg_z"),
                              s.vector("dual_slack"))
            assert r[0] <= 2 * gap * 0.1 and r[1] <= 2 * gap and r[2] <= 2 * gap, r
        gap = max(gap * 0.98, 1e-4)


def test_fp32_polish_gives_exact_fp64_bounds():
    """Bound-active coordinates of the FP32 projection carry the FP64 bound
    exactly, so D_c·x̂ reproduces the original bound bit for bit."""
    m = G.set_cover(133, 201, 3)
    xt = np.zeros(201)
    pp, s, status = project_with(P, P.PDFPState, m, xt, tol=1e-6, precision="fp32")
    x_hat = s.pdhg_x
    at_lo = x_hat == pp.var_lb
    at_hi = x_hat == pp.var_ub
    assert at_lo.sum() > 50                      # most columns stay at 0
    x = pp.D_c * x_hat
    assert np.all(x[at_lo] == 0.0)
    assert np.all(np.abs(x[at_hi] - 1.0) <= 1e-15)
    assert max_violation(m, x) <= 1e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_edge_cases(precision):
    """No rows (projection = clamp), an empty row and an empty column,
    infinite bounds, and the iteration limit."""
    say("edges", precision)
    import scipy.sparse as sp
    n = 5
    milp = G.MILP(c=np.zeros(n), c0=0.0, A=sp.csc_matrix((0, n)), b=np.zeros(0),
                  row_sense=np.zeros(0, dtype="<U2"), l=np.zeros(n), u=np.ones(n),
                  int_idx=np.arange(n))
    xt = np.array([-1.0, 0.3, 2.0, 1.0, 0.0])
    x, st = P.project(milp, xt, tol=1e-7, precision=precision)
    assert st == "OPTIMAL"
    np.testing.assert_allclose(x, np.clip(xt, 0, 1), atol=1e-5)
    # an empty row (0 ≤ 1) and an empty column (x4 free in [0, inf))
    A = sp.csc_matrix(np.array([[1.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.0, 1.0, 1.0, 0.0]]))
    milp = G.MILP(c=np.zeros(4), c0=0.0, A=A, b=np.array([1.0, 1.0, 1.0]),
                  row_sense=np.array(["le", "le", "eq"]), l=np.zeros(4),
                  u=np.array([1.0, 1.0, 1.0, np.inf]), int_idx=np.array([0, 1, 2]))
    xt = np.array([1.0, 1.0, 1.0, 5.0])
    pp, s, st = project_with(P, P.PDFPState, milp, xt, tol=1e-7, precision=precision)
    pq, o, so = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-7)
    assert st == so == "OPTIMAL"
    assert abs(s.info().l1_distance - o.info().l1_distance) <= 1e-5
    m = G.set_cover(133, 201, 1)
    pp, s, st = project_with(P, P.PDFPState, m, np.zeros(201), tol=1e-14, iteration_limit=400,
                             precision=precision)
    assert st == "ITERATION_LIMIT" and s.total_count == 400


@pytest.mark.parametrize("name,maker", [
    ("mixed_2000x4000", lambda: G.mixed(2000, 4000, 1)),
    ("mixed_50000x100000", lambda: G.mixed(50_000, 100_000, 1)),
])
def test_block_graph_replays_equal_stream_launches(name, maker):
    """The FP32 general passes replayed as one captured graph per block give
    the same iterates, bit for bit, as the same launches on the stream: a
    warm-started sequence of projections (many replays, restarts, host reads
    and uploads between them), each run on a fresh state with the same
    schedule (the schedule sets each row's summation order)."""
    say("block graph", name)
    m = maker()
    pp = projector_problem(m)
    n = m.A.shape[1]
    targets = [G.binary_targets(m, 90 + k) / pp.D_c for k in range(3)]
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-5,
                         iteration_limit=100_000)
    runs = []
    for graph in (False, True, True):
        s = hip_state(pp, "fp32")
        s.set_kernel_path("general")
        s.set_schedule(512, 512, 0, 0)
        s.set_block_graph(graph)
        s.set_x(np.zeros(n))
        seq = []
        for xt in targets:
            P.set_target(s, xt)
            assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
            info = s.info()
            assert info.block_graph == int(graph)
            assert (info.npb_a, info.npb_at, info.persist_a, info.persist_at) == (512, 512, 0, 0)
            seq.append((info.total_count, info.n_restarts, info.l1_distance,
                        s.vector("pdhg_x"), s.vector("pdhg_z")))
        s.close()
        runs.append(seq)
    assert sum(r[0] for r in runs[1]) > 3 * 200  # several replays per projection
    for seq in runs[1:]:
        for a, b in zip(runs[0], seq):
            assert a[:3] == b[:3]
            assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("name,maker,path", [
    ("p0201", lambda: read_mps(example_path("p0201")), "auto"),
    ("mixed_2000x4000", lambda: G.mixed(2000, 4000, 1), "general")])
def test_target_in_original_units_equals_host_division(precision, name, maker, path):
    """set_target_original (x̃ in original units, x̃ ./ D_c on the device, the
    pump's src/pump.jl:331) gives the projections of set_target on the
    host's x̃ ./ D_c bit for bit: the stored target, the iteration counts,
    the ℓ1 gap (whose denominator is summed from the target), the points;
    a warm-started sequence, each form on a fresh state with the same
    schedule."""
    say("target original units", precision, name, path)
    m = maker()
    pp = projector_problem(m)
    n = m.A.shape[1]
    targets = [G.binary_targets(m, 30 + k) for k in range(3)]
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=100_000)
    runs = []
    for device_division in (False, True):
        s = hip_state(pp, precision)
        s.set_kernel_path(path)
        s.set_schedule(512, 512, 0, 0)
        s.set_x(np.zeros(n))
        if device_division:
            P.set_column_scale(s, pp.D_c)
        seq = []
        for xt in targets:
            if device_division:
                P.set_target_original(s, xt)
            else:
                P.set_target(s, xt / pp.D_c)
            st = P.pdfp_solve(s, p, 1.0)
            info = s.info()
            seq.append((st, info.total_count, info.relative_l1_gap, info.l1_distance,
                        s.vector("x_tilde"), s.pdhg_x.copy()))
        s.close()
        runs.append(seq)
    for a, b in zip(*runs):
        assert a[0] == "OPTIMAL" and a[:4] == b[:4], (a[:4], b[:4])
        assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


def test_target_in_original_units_errors():
    """No column scaling yet: a state error; a zero or short scaling: refused."""
    m = G.set_cover(133, 201, 1)
    pp = projector_problem(m)
    s = hip_state(pp, "fp32")
    try:
        xt = G.binary_targets(m, 1)
        with pytest.raises(LibraryError, match="column scaling"):
            P.set_target_original(s, xt)
        bad = pp.D_c.copy()
        bad[3] = 0.0
        with pytest.raises(LibraryError, match="positive"):
            P.set_column_scale(s, bad)
        with pytest.raises(ValueError):
            P.set_column_scale(s, pp.D_c[:-1])
        P.set_column_scale(s, pp.D_c)
        P.set_target_original(s, xt)
        assert np.array_equal(s.vector("x_tilde"),
                              (xt / pp.D_c).astype(np.float32).astype(np.float64))
    finally:
        s.close()
