"""The FP32 form's one-workgroup blocks (whole evaluation blocks in one
launch for instances whose operators fit in LDS) against the general passes
and the oracle.

Both paths run the same iteration (src/pdfp.jl:119-163) in FP32 with FP64
sums; only the order of the sums inside a row differs, so they agree to FP32
rounding, not bit for bit: the same termination, iteration counts within a
few evaluation blocks, and ℓ1 distances within 1e-5 relative.  Every block's
KKT scalars are FP64 evaluations of the returned iterates, checked by the
oracle's compute_residual!.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd import pdfp as P
from pdfp_amd.milp import read_mps
from pdfp_amd.pdfp import SolverParameters, projector_problem

from .conftest import example_path

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve(pp, xt, path, tol, freq=200, pw=1.0):
    s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                    precision="fp32")
    s.set_kernel_path(path)
    P.set_target(s, xt / pp.D_c)
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000, termination_evaluation_frequency=freq)
    st = P.pdfp_solve(s, p, pw)
    return s, st


def instances():
    yield "set_cover_133x201", G.set_cover(133, 201, 1), None
    yield "p0201", read_mps(example_path("p0201")), None
    yield "pk1", read_mps(example_path("pk1")), None
    yield "toy2", G.toy2(), np.array([3.0, 0.0])


@pytest.mark.parametrize("name,milp,xt", list(instances()), ids=lambda v: v if isinstance(v, str) else "")
def test_small_blocks_match_general_passes(name, milp, xt):
    say("small vs general", name)
    pp = projector_problem(milp)
    if xt is None:
        xt = G.binary_targets(milp, 7)
    tol = 1e-5
    s1, st1 = solve(pp, xt, "auto", tol)
    s2, st2 = solve(pp, xt, "general", tol)
    assert s1.info().small_path == 1 and s2.info().small_path == 0
    assert st1 == st2 == "OPTIMAL"
    assert abs(s1.total_count - s2.total_count) <= 5 * 200, (s1.total_count, s2.total_count)
    l1a, l1b = s1.info().l1_distance, s2.info().l1_distance
    assert abs(l1a - l1b) <= 1e-5 * max(1.0, abs(l1b)), (l1a, l1b)
    o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    O.set_target(o, xt / pp.D_c)
    o.initialize(None, 1.0)
    r = o.residual_of(s1.pdhg_x, s1.vector("pdhg_y"), s1.vector("pdhg_z"),
                      s1.vector("dual_slack"))
    assert max(r[0], r[1], r[2]) <= 10 * tol, r


def test_small_blocks_restart_and_every_iteration_evaluation():
    """Evaluation every iteration (the first step of a block is then also its
    last) and restarts (the evaluation after the first step of a block) take
    the one-workgroup kernel's two special cases."""
    say("small freq=1")
    milp = G.set_cover(133, 201, 2)
    pp = projector_problem(milp)
    xt = G.binary_targets(milp, 11)
    s1, st1 = solve(pp, xt, "auto", 1e-4, freq=1)
    s2, st2 = solve(pp, xt, "general", 1e-4, freq=1)
    assert st1 == st2 == "OPTIMAL"
    assert s1.info().small_path == 1
    assert abs(s1.total_count - s2.total_count) <= max(20, s2.total_count // 20)
    s3, st3 = solve(pp, xt, "auto", 1e-6, freq=64)
    assert st3 == "OPTIMAL" and s3.info().n_restarts >= 1


def test_small_blocks_warm_sequence_and_mixed_rows():
    """A pump-like sequence of warm-started projections on a mixed model
    with the primal residual in original units (src/pump.jl:304)."""
    say("small warm sequence")
    milp = read_mps(example_path("pk1"))
    pp = projector_problem(milp)
    states = []
    for path in ("auto", "general"):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        s.set_kernel_path(path)
        P.original_primal_residual(s, pp.D_r)
        s.set_x(np.clip(np.zeros(pp.A.shape[1]), milp.l, milp.u) / pp.D_c)
        states.append(s)
    gap = 1e-2
    for k in range(4):
        xt_orig = G.binary_targets(milp, 50 + k)
        xt = xt_orig / pp.D_c
        den = 1.0 + np.abs(xt_orig[milp.int_idx]).sum()   # the gap's denominator
        p = SolverParameters(eps_optimal_relative=gap, eps_feasible_relative=gap * 0.1,
                             iteration_limit=200_000)
        res = []
        for s in states:
            P.set_target(s, xt)
            res.append((P.pdfp_solve(s, p, 1.0), s.info().l1_distance))
        assert res[0][0] == res[1][0] == "OPTIMAL"
        assert states[0].info().small_path == 1
        assert abs(res[0][1] - res[1][1]) <= 4 * gap * den, (res, den)
        gap *= 0.5
