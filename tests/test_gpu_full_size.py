"""GPU tests at the benchmark's full sizes: the 1M-nonzero mixed MILP
(50k × 100k), the 5M-nonzero power-law one (500k × 1M), the 4M-nonzero
general-integer one with 8 batched trajectories (200k × 400k) and the pack of
64 uneven set covers (15.5M nonzeros).

The oracle is too slow to run these projections, so the checks are the
size-independent ones: every FP32 projection (single and batched) is
re-evaluated in FP64 by the oracle's compute_residual! (src/pdfp.jl:200-218)
and meets the tolerance (×10 for FP32 rounding); and its ℓ1 distance agrees
with the FP64 form's projection of the same target within what the gap
tolerance allows (both are within ε·(1 + Σ_I w|x̃|) of the optimum, so they
differ by at most 2ε·(1 + Σ_I w|x̃|)).  The FP64 form's own projections are
re-evaluated the same way: they meet the tolerance, and the residuals the
device reported are the oracle's to 1e-9 (of the tolerance, for residuals
far below it)."""
import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd import pdfp as P
from pdfp_amd.batch import PDFPBatch
from pdfp_amd.pdfp import SolverParameters, projector_problem

pytestmark = pytest.mark.gpu

TOL = 1e-4

WORKLOADS = {
    "mixed_50000x100000_1M": lambda: G.mixed(50_000, 100_000, 1),
    "powerlaw_500000x1000000_5M": lambda: G.mixed(500_000, 1_000_000, 1, power_law=True,
                                                  nnz=5_000_000, max_len=5000),
}


class _Kkt:
    """The oracle's FP64 residuals of a given point; the power iteration is
    cut to one step, since the residuals do not depend on the step size."""

    def __init__(self, pp):
        self.o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
        self.p = SolverParameters(sv_max_iterations=1)

    def __call__(self, xt_scaled, x, y, z, ds):
        O.set_target(self.o, xt_scaled)
        self.o.initialize(self.p, 1.0)
        return self.o.residual_of(x, y, z, ds)


def _denominator(pp, xt_scaled):
    ii = np.asarray(pp.is_int, dtype=bool)
    return 1.0 + float(np.sum(np.asarray(pp.w)[ii] * np.abs(xt_scaled[ii])))


def _fp64_against_oracle(kkt, s64, xt_scaled, tag):
    """The FP64 form's own projection, re-evaluated by the oracle: it meets
    the tolerance, and the device's residuals and ℓ1 distance are the
    oracle's evaluation of the same point to 1e-9 of the larger of the value
    and the tolerance (they differ only in summation order: a residual far
    below the tolerance is a difference of O(1) norms, and so carries their
    rounding, about 1e-15, as an absolute error).  Returns the ℓ1 distance."""
    info = s64.info()
    r = kkt(xt_scaled, s64.pdhg_x, s64.vector("pdhg_y"), s64.vector("pdhg_z"),
            s64.vector("dual_slack"))
    assert max(r[0], r[1], r[2]) <= TOL * (1 + 1e-6), (tag, r)
    for dev, orc in ((info.relative_primal_residual, r[0]),
                     (info.relative_dual_residual, r[1]), (info.relative_l1_gap, r[2]),
                     (info.l1_distance, r[3])):
        assert abs(dev - orc) <= 1e-9 * max(abs(orc), TOL), (tag, dev, orc)
    return info.l1_distance


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_projections_fp32_single_and_batch_against_fp64(name):
    print("full size", name, flush=True)
    milp = WORKLOADS[name]()
    pp = projector_problem(milp)
    n = milp.A.shape[1]
    x_start = np.clip(np.zeros(n), milp.l, milp.u) / pp.D_c
    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    p = SolverParameters(eps_optimal_relative=TOL, eps_feasible_relative=TOL,
                         iteration_limit=200_000)
    kkt = _Kkt(pp)
    targets = [G.binary_targets(milp, 70 + k) / pp.D_c for k in range(8)]

    # FP64 form (the reference's arithmetic): ℓ1 of the first two targets,
    # warm-started one after the other as the pump does.
    s64 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                      precision="fp64")
    s64.set_x(x_start)
    l1_64 = []
    for k, xt in enumerate(targets[:2]):
        P.set_target(s64, xt)
        assert P.pdfp_solve(s64, p, pw) == "OPTIMAL"
        l1_64.append(_fp64_against_oracle(kkt, s64, xt, ("fp64", name, k)))
    s64.close()

    # Single FP32 form, the same sequence.
    s32 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                      precision="fp32")
    s32.set_x(x_start)
    for k, xt in enumerate(targets[:2]):
        P.set_target(s32, xt)
        assert P.pdfp_solve(s32, p, pw) == "OPTIMAL"
        r = kkt(xt, s32.pdhg_x, s32.vector("pdhg_y"), s32.vector("pdhg_z"),
                s32.vector("dual_slack"))
        assert r[0] <= 10 * TOL and r[1] <= 10 * TOL and r[2] <= 10 * TOL, ("fp32", k, r)
        # r[3]: the ℓ1 distance of the returned point, evaluated in FP64
        assert abs(r[3] - l1_64[k]) <= 2 * TOL * _denominator(pp, xt), ("fp32", k, r[3], l1_64[k])
    s32.close()

    # The 8-trajectory batch, one target per trajectory.
    bt = PDFPBatch(8, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    for k in range(8):
        bt.set_x(k, x_start)
        bt.set_target(k, targets[k])
    assert bt.solve(p, pw) == ["OPTIMAL"] * 8
    for k in range(8):
        r = kkt(targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"), bt.vector(k, "pdhg_z"),
                bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * TOL and r[1] <= 10 * TOL and r[2] <= 10 * TOL, ("batch", k, r)
        if k < 2:
            assert abs(r[3] - l1_64[k]) <= 2 * TOL * _denominator(pp, targets[k]), \
                ("batch", k, r[3], l1_64[k])
    bt.close()


def test_full_size_general_integer_batch_of_8():
    """The general-integer, equality-heavy configuration at full size:
    200,000 rows (40 % equalities) × 400,000 integer columns in [0, U_j],
    4M coefficients in [-9, 9], 8 trajectories on one batch.  Every
    trajectory's point meets the tolerance under the oracle's FP64 residuals,
    and the first two trajectories' ℓ1 distances agree with the FP64 form's
    projections of the same targets within the gap tolerance."""
    print("full size general integer batch", flush=True)
    milp = G.general_integer(200_000, 400_000, 1)
    pp = projector_problem(milp)
    n = milp.A.shape[1]
    x_start = np.clip(np.zeros(n), milp.l, milp.u) / pp.D_c
    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    p = SolverParameters(eps_optimal_relative=TOL, eps_feasible_relative=TOL,
                         iteration_limit=200_000)
    kkt = _Kkt(pp)
    targets = [G.binary_targets(milp, 90 + k) / pp.D_c for k in range(8)]
    s64 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                      precision="fp64")
    l1_64 = []
    for k, xt in enumerate(targets[:2]):
        s64.set_x(x_start)
        P.set_target(s64, xt)
        assert P.pdfp_solve(s64, p, pw) == "OPTIMAL"
        l1_64.append(_fp64_against_oracle(kkt, s64, xt, ("fp64 general integer", k)))
    s64.close()
    bt = PDFPBatch(8, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    for k in range(8):
        bt.set_x(k, x_start)
        bt.set_target(k, targets[k])
    assert bt.solve(p, pw) == ["OPTIMAL"] * 8
    for k in range(8):
        r = kkt(targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"), bt.vector(k, "pdhg_z"),
                bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * TOL and r[1] <= 10 * TOL and r[2] <= 10 * TOL, ("batch", k, r)
        if k < 2:
            assert abs(r[3] - l1_64[k]) <= 2 * TOL * _denominator(pp, targets[k]), \
                ("batch", k, r[3], l1_64[k])
    bt.close()


def test_full_size_pack_of_64_uneven_set_covers():
    """The 64-instance configuration at full size: set covers of uneven sizes
    (seeds 1..64, rows e^U[ln 2k, ln 40k], columns e^U[ln 4k, ln 80k]; 15.5M
    nonzeros in all) in one block-diagonal pack.  Every instance's point
    meets the tolerance under the oracle's FP64 residuals, and its ℓ1
    distance agrees with the single FP64 form's projection of the same
    target within the gap tolerance."""
    from pdfp_amd.pack import PDFPPack

    print("full size pack of 64", flush=True)
    milps = G.uneven_set_covers(64)
    pps = [projector_problem(m) for m in milps]
    assert sum(m.A.nnz for m in milps) > 15_000_000
    p = SolverParameters(eps_optimal_relative=TOL, eps_feasible_relative=TOL,
                         iteration_limit=200_000)
    pws = [float(np.sqrt(1.0 + len(m.int_idx) / max(len(m.b), 1))) for m in milps]
    targets = [G.binary_targets(m, 60 + k) / pp.D_c for k, (m, pp) in enumerate(zip(milps, pps))]
    pk = PDFPPack(pps)
    for k, t in enumerate(targets):
        pk.set_target(k, t)
    assert pk.solve(p, pws) == ["OPTIMAL"] * 64
    for k, pp in enumerate(pps):
        kkt = _Kkt(pp)
        r = kkt(targets[k], pk.pdhg_x(k), pk.vector(k, "pdhg_y"), pk.vector(k, "pdhg_z"),
                pk.vector(k, "dual_slack"))
        assert r[0] <= 10 * TOL and r[1] <= 10 * TOL and r[2] <= 10 * TOL, ("pack", k, r)
        s64 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                          precision="fp64")
        P.set_target(s64, targets[k])
        assert P.pdfp_solve(s64, p, pws[k]) == "OPTIMAL"
        l1_64 = _fp64_against_oracle(kkt, s64, targets[k], ("fp64 pack", k))
        s64.close()
        assert abs(r[3] - l1_64) <= 2 * TOL * _denominator(pp, targets[k]), \
            ("pack", k, r[3], l1_64)
    pk.close()


def test_full_size_powerlaw_batch_of_16():
    """The 5M-nonzero power-law instance with 16 trajectories on one batch
    (batch width 16): every trajectory's point meets the tolerance under the
    oracle's FP64 residuals, and two of them agree in ℓ1 with the FP64 form's
    projections of the same targets within the gap tolerance."""
    print("full size powerlaw batch of 16", flush=True)
    milp = WORKLOADS["powerlaw_500000x1000000_5M"]()
    pp = projector_problem(milp)
    n = milp.A.shape[1]
    x_start = np.clip(np.zeros(n), milp.l, milp.u) / pp.D_c
    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    p = SolverParameters(eps_optimal_relative=TOL, eps_feasible_relative=TOL,
                         iteration_limit=200_000)
    kkt = _Kkt(pp)
    targets = [G.binary_targets(milp, 170 + k) / pp.D_c for k in range(16)]
    s64 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                      precision="fp64")
    l1_64 = {}
    for k in (0, 15):
        s64.set_x(x_start)
        P.set_target(s64, targets[k])
        assert P.pdfp_solve(s64, p, pw) == "OPTIMAL"
        l1_64[k] = _fp64_against_oracle(kkt, s64, targets[k], ("fp64 powerlaw", k))
    s64.close()
    bt = PDFPBatch(16, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    for k in range(16):
        bt.set_x(k, x_start)
        bt.set_target(k, targets[k])
    assert bt.solve(p, pw) == ["OPTIMAL"] * 16
    for k in range(16):
        r = kkt(targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"), bt.vector(k, "pdhg_z"),
                bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * TOL and r[1] <= 10 * TOL and r[2] <= 10 * TOL, ("batch16", k, r)
        if k in l1_64:
            assert abs(r[3] - l1_64[k]) <= 2 * TOL * _denominator(pp, targets[k]), \
                ("batch16", k, r[3], l1_64[k])
    bt.close()
