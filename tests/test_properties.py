"""Property tests of the oracle and the host rounder on random tiny MILPs
(tests/properties.py): the oracle's ℓ1 projection (src/pdfp.jl:252-288) at
tolerance 1e-8 ends OPTIMAL at a point that satisfies the rows and bounds
and whose ℓ1 distance is HiGHS's optimum of the same projection LP; the C++
rounder equals the Python restatement of src/rounding.jl bit for bit; the
oracle's root LP (src/pdlp.jl:381-393) at tolerance 1e-7, with a random
objective, ends OPTIMAL at HiGHS's LP optimum."""
import numpy as np
from hypothesis import HealthCheck, assume, given, settings

from oracle import oracle as O
from pdfp_amd.milp import max_violation
from pdfp_amd.rounding import Rounder

from .golden.make_fixtures import l1_projection_value, lp_value
from .helpers import l1_distance, project_with
from .properties import fractional_point, instances, integer_target, tiny_milp

SETTINGS = settings(max_examples=100, deadline=None, derandomize=True,
                    suppress_health_check=[HealthCheck.too_slow])


@SETTINGS
@given(instances)
def test_oracle_projection_is_highs_optimum(spec):
    m, n, density, eq_frac, int_frac, seed = spec
    milp = tiny_milp(m, n, density, eq_frac, int_frac, seed)
    xt = integer_target(milp, seed)
    pp, s, status = project_with(O, O.OraclePDFPState, milp, xt, tol=1e-8)
    assert status == "OPTIMAL"
    x = s.pdhg_x * pp.D_c
    ref = l1_projection_value(milp, xt)
    assert abs(l1_distance(milp, x, xt) - ref) <= 1e-5 * max(1.0, ref), (spec, ref)
    scale = 1.0 + np.abs(milp.b).max()
    assert max_violation(milp, x) <= 1e-6 * scale, spec
    assert np.all(x >= milp.l - 1e-9) and np.all(x <= milp.u + 1e-9)


@SETTINGS
@given(instances)
def test_native_rounder_equals_python_restatement(spec):
    m, n, density, eq_frac, int_frac, seed = spec
    milp = tiny_milp(m, n, density, eq_frac, int_frac, seed)
    assume(len(milp.int_idx) > 0)
    x = fractional_point(milp, seed)
    for schedule in ("all", "regroup"):
        outs = []
        for backend in ("native", "python"):
            r = Rounder(milp, backend=backend)
            r.schedule = schedule
            rng = np.random.Generator(np.random.PCG64(seed))
            outs.append((r.round_guarded(x, rng, 12), r.round_sequential(x, rng, 12)))
        for a, b in zip(*outs):
            assert np.array_equal(a, b), (spec, schedule)


@SETTINGS
@given(instances)
def test_oracle_root_lp_is_highs_optimum(spec):
    """Without cuPDLPx termination (the pump's root LP on pure integer
    models, src/pump.jl:269-271) the oracle always reaches HiGHS's optimum.
    With it (mixed models) the ray test of src/pdlp.jl:287-288 may call a
    feasible LP PRIMAL_INFEASIBLE when every variable is bounded: the dual
    ray's infeasibility (src/pdlp.jl:270-278) is then 0 by construction, so
    any positive dual ray objective passes it.  The restatement keeps that;
    the first such draw is 3 × 3 (spec (3, 3, 0.625, 0.25, 1.0, 3))."""
    m, n, density, eq_frac, int_frac, seed = spec
    milp = tiny_milp(m, n, density, eq_frac, int_frac, seed)
    milp.c = np.random.default_rng(seed + 3).standard_normal(n)
    ref = lp_value(milp)
    for cupdlpx in (False, True):
        x, ds, status, it = O.solve_lp_relaxation(milp, tolerance=1e-7,
                                                  cupdlpx_termination=cupdlpx)
        if cupdlpx and status == "PRIMAL_INFEASIBLE":
            assert np.all(np.isfinite(milp.l)) and np.all(np.isfinite(milp.u))
            continue
        assert status == "OPTIMAL", (spec, cupdlpx, status)
        obj = float(milp.c @ x)
        assert abs(obj - ref) <= 1e-5 * (1.0 + abs(ref)), (spec, cupdlpx, obj, ref)
        assert max_violation(milp, x) <= 1e-5 * (1.0 + np.abs(milp.b).max()), spec


def test_cupdlpx_ray_test_calls_a_bounded_feasible_lp_infeasible():
    """The draw above, pinned: feasible (HiGHS's optimum 2.1062...), OPTIMAL
    without cuPDLPx termination, PRIMAL_INFEASIBLE with it after 200
    iterations (src/pdlp.jl:355-358, the first evaluation)."""
    milp = tiny_milp(3, 3, 0.625, 0.25, 1.0, 3)
    milp.c = np.random.default_rng(6).standard_normal(3)
    assert abs(lp_value(milp) - 2.1062315089735164) <= 1e-9
    x, ds, st, it = O.solve_lp_relaxation(milp, tolerance=1e-7, cupdlpx_termination=False)
    assert st == "OPTIMAL"
    x, ds, st, it = O.solve_lp_relaxation(milp, tolerance=1e-7, cupdlpx_termination=True)
    assert (st, it) == ("PRIMAL_INFEASIBLE", 200)
