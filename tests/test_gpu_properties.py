"""Device parity on random tiny MILPs (tests/properties.py), drawn by
hypothesis with a fixed derandomized sequence: the FP64 form, on both
kernel paths, follows the oracle's trajectory (same termination and
iteration count, pdhg x within 1e-9); the FP32 form, at the pump's
tolerance 1e-4 (FP32 iterates can sit above 1e-5 on such instances: one
drawn instance reached ITERATION_LIMIT at 1e-5), ends OPTIMAL at a point
whose KKT residuals, evaluated in FP64 by the oracle, are within 10× the
tolerance and whose ℓ1 distance is the oracle's optimum to the tolerance; the
batch of 8 trajectories on the general passes reaches the single FP32
projector's ℓ1 distances; the root LP follows the oracle's trajectory (same
termination and iteration count) on both kernel paths; the device scaling
and the device feasibility check equal the oracle's and the host's bit for
bit; a pack of 2-6 drawn instances equals the single FP32 projector bit for
bit on the one-workgroup path and reaches its distances on the general
passes; the batch with its A pass forced over the column slices (every row,
or the rows longer than 2; drawn instances have fewer columns than slices,
empty slices, rows and columns) reaches the same distances."""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from oracle import oracle as O
from pdfp_amd import pdfp as P
from pdfp_amd import pdlp as L
from pdfp_amd.batch import PDFPBatch
from pdfp_amd.pack import PDFPPack
from pdfp_amd.milp import DeviceViolationCheck, max_violation
from pdfp_amd.pdfp import SolverParameters, projector_problem
from pdfp_amd.preconditioner import scale_problem

from .helpers import l1_distance, project_with
from .properties import instances, integer_target, tiny_milp

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=40, deadline=None, derandomize=True,
                    suppress_health_check=[HealthCheck.too_slow])


def _state(path, precision):
    def make(*a, **kw):
        s = P.PDFPState(*a, precision=precision, **kw)
        s.set_kernel_path(path)
        return s
    return make


@pytest.mark.parametrize("path", ["auto", "one-workgroup"])
@SETTINGS
@given(spec=instances)
def test_fp64_follows_the_oracle_on_random_instances(path, spec):
    milp = tiny_milp(*spec)
    xt = integer_target(milp, spec[-1])
    print("fp64 property", path, spec, flush=True)
    pp, s, st = project_with(P, _state(path, "fp64"), milp, xt, tol=1e-7)
    pq, o, so = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-7)
    try:
        assert st == so == "OPTIMAL", (st, so)
        assert s.total_count == o.total_count, (s.total_count, o.total_count)
        np.testing.assert_allclose(s.pdhg_x, o.pdhg_x, rtol=0, atol=1e-9)
    finally:
        s.close()


@SETTINGS
@given(spec=instances)
def test_fp32_reaches_the_oracle_distance_on_random_instances(spec):
    milp = tiny_milp(*spec)
    xt = integer_target(milp, spec[-1])
    print("fp32 property", spec, flush=True)
    tol = 1e-4
    pp, s, st = project_with(P, _state("auto", "fp32"), milp, xt, tol=tol)
    pq, o, so = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-8)
    try:
        assert st == so == "OPTIMAL", (st, so)
        r = o.residual_of(s.pdhg_x, s.vector("pdhg_y"), s.vector("pdhg_z"),
                          s.vector("dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, r
        den = 1.0 + np.abs(xt[milp.int_idx]).sum()
        lo = l1_distance(milp, o.pdhg_x * pq.D_c, xt)
        assert abs(l1_distance(milp, s.pdhg_x * pp.D_c, xt) - lo) <= 10 * tol * den
    finally:
        s.close()


@settings(max_examples=15, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow])
@given(spec=instances)
def test_batch_matches_single_fp32_on_random_instances(spec):
    milp = tiny_milp(*spec)
    pp = projector_problem(milp)
    print("batch property", spec, flush=True)
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=200_000)
    targets = [integer_target(milp, spec[-1] + 31 * k) / pp.D_c for k in range(8)]
    ref = []
    for k in range(8):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        ref.append(s.info().l1_distance)
        s.close()
    bt = PDFPBatch(8, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    try:
        bt.set_kernel_path("general")
        for k in range(8):
            bt.set_target(k, targets[k])
        assert bt.solve(p, 1.0) == ["OPTIMAL"] * 8
        for k in range(8):
            l1b = bt.info(k).l1_distance
            assert abs(l1b - ref[k]) <= 1e-3 * max(1.0, ref[k]), (k, l1b, ref[k])
    finally:
        bt.close()


@pytest.mark.parametrize("cupdlpx", [False, True])
@pytest.mark.parametrize("path", ["auto", "general"])
@SETTINGS
@given(spec=instances)
def test_root_lp_follows_the_oracle_on_random_instances(path, cupdlpx, spec):
    """Both terminations, whatever the oracle concludes (under cuPDLPx
    termination that includes the reference's PRIMAL_INFEASIBLE verdicts on
    bounded feasible LPs, tests/test_properties.py)."""
    milp = tiny_milp(*spec)
    milp.c = np.random.default_rng(spec[-1] + 3).standard_normal(len(milp.c))
    print("root lp property", path, cupdlpx, spec, flush=True)
    xo, dso, so, ito = O.solve_lp_relaxation(milp, tolerance=1e-6, cupdlpx_termination=cupdlpx)
    x, ds, st, it = L.solve_lp_relaxation(milp, tolerance=1e-6, time_limit=30.0,
                                          cupdlpx_termination=cupdlpx, kernel_path=path)
    assert st == so, (st, so)
    assert it == ito, (it, ito)
    if so == "OPTIMAL":
        np.testing.assert_allclose(x, xo, rtol=0, atol=1e-7 * (1.0 + np.abs(xo).max()))


@SETTINGS
@given(spec=instances)
def test_scaling_and_check_equal_the_host_on_random_instances(spec):
    milp = tiny_milp(*spec)
    print("scale/check property", spec, flush=True)
    con_lb = np.where(milp.row_sense == "eq", milp.b, -np.inf)
    a = scale_problem(milp.A, milp.c, milp.l, milp.u, con_lb, milp.b, backend="device")
    b = O.scale_problem(milp.A, milp.c, milp.l, milp.u, con_lb, milp.b)
    for k in ("D_r", "D_c"):
        assert np.array_equal(getattr(a, k), b[k]), k
    assert np.array_equal(a.A.data, b["A"].data)
    chk = DeviceViolationCheck(milp)
    rng = np.random.default_rng(spec[-1])
    for _ in range(3):
        x = rng.random(len(milp.c)) * milp.u
        assert chk(x) == max_violation(milp, x)


@settings(max_examples=15, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow])
@given(specs=st.lists(instances, min_size=2, max_size=6))
def test_pack_equals_single_projectors_on_random_instances(specs):
    milps = [tiny_milp(*sp_) for sp_ in specs]
    pps = [projector_problem(m) for m in milps]
    print("pack property", specs, flush=True)
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=200_000)
    targets = [integer_target(m, sp_[-1] + 17) / pp.D_c for m, sp_, pp in zip(milps, specs, pps)]
    ref = []
    for pp, t in zip(pps, targets):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, t)
        stt = P.pdfp_solve(s, p, 1.0)
        ref.append((stt, s.total_count, s.pdhg_x.copy(), s.info().l1_distance,
                    s.info().small_path))
        s.close()
    for path in ("auto", "general"):
        pk = PDFPPack(pps)
        try:
            pk.set_kernel_path(path)
            for k, t in enumerate(targets):
                pk.set_target(k, t)
            terms = pk.solve(p, 1.0)
            for k, r in enumerate(ref):
                if path == "auto" and r[4] == 1:
                    assert terms[k] == r[0] and pk.total_count(k) == r[1], (k, terms[k], r[:2])
                    assert np.array_equal(pk.pdhg_x(k), r[2]), k
                elif r[0] == "OPTIMAL":
                    assert terms[k] == "OPTIMAL", (k, path)
                    l1 = pk.info(k).l1_distance
                    assert abs(l1 - r[3]) <= 1e-3 * max(1.0, r[3]), (k, path, l1, r[3])
        finally:
            pk.close()


@pytest.mark.parametrize("min_len", [0, 2])
@settings(max_examples=15, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow])
@given(spec=instances)
def test_batch_column_slices_on_random_instances(min_len, spec):
    milp = tiny_milp(*spec)
    pp = projector_problem(milp)
    print("batch slices property", min_len, spec, flush=True)
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=200_000)
    targets = [integer_target(milp, spec[-1] + 13 * k) / pp.D_c for k in range(8)]
    ref = []
    for k in range(8):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        ref.append(s.info().l1_distance)
        s.close()
    bt = PDFPBatch(8, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    try:
        bt.set_kernel_path("general")
        if pp.A.nnz == 0:
            return  # nothing to slice (the library refuses, as documented)
        bt.set_column_slices(True)
        bt.set_slice_min_len(min_len)
        for k in range(8):
            bt.set_target(k, targets[k])
        assert bt.solve(p, 1.0) == ["OPTIMAL"] * 8
        assert bt.info(0).column_slices > 0 and bt.info(0).slice_min_len == min_len
        for k in range(8):
            l1b = bt.info(k).l1_distance
            assert abs(l1b - ref[k]) <= 1e-3 * max(1.0, ref[k]), (k, l1b, ref[k])
    finally:
        bt.close()
