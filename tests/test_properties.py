"""Property tests of the oracle and the host rounder on random tiny MILPs
(tests/properties.py): the oracle's ℓ1 projection (src/pdfp.jl:252-288) at
tolerance 1e-8 ends OPTIMAL at a point that satisfies the rows and bounds
and whose ℓ1 distance is HiGHS's optimum of the same projection LP; the C++
rounder equals the Python restatement of src/rounding.jl bit for bit."""
import numpy as np
from hypothesis import HealthCheck, assume, given, settings

from oracle import oracle as O
from pdfp_amd.milp import max_violation
from pdfp_amd.rounding import Rounder

from .golden.make_fixtures import l1_projection_value
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
