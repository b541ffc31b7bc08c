"""CPU tests of the host side of the pump: propagation, guarded CPR rounding
and the pump loop (src/propagation.jl, src/rounding.jl, src/pump.jl).  The
projector and the root LP are the oracle here (the device runs them in the
GPU tests); the pump code under test is the product's."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from pdfp_amd import generators as G
from pdfp_amd.milp import MILP, is_feasible, read_mps
from pdfp_amd.propagation import PropagationData, PropagationState, activities
from pdfp_amd.pump import feasibility_pump, flip_farthest, round_to_nearest
from pdfp_amd.rounding import Rounder

from .conftest import example_path
from .helpers import oracle_projector_module, oracle_root_lp


def pump(milp, **kw):
    return feasibility_pump(milp, projector=oracle_projector_module(), root_lp=oracle_root_lp,
                            scaling="host", **kw)


@pytest.mark.parametrize("name", ["p0201", "gen-ip054", "pk1", "markshare2"])
def test_pump_on_examples(name):
    """test/runtests.jl:73-79."""
    m = read_mps(example_path(name))
    r = pump(m, time_limit=120.0)
    assert r.found
    assert is_feasible(m, r.x)
    assert abs(r.objective - (float(m.c @ r.x) + m.c0)) <= 1e-9 * (1 + abs(r.objective))
    assert r.time >= r.root_lp_time and r.pdhg_iterations >= 0


def test_pump_on_set_cover():
    m = G.set_cover(133, 201, 1)
    r = pump(m, time_limit=60.0, batch_policy="always")
    assert r.found and is_feasible(m, r.x)
    assert r.iterations >= 0 and r.log and "FEASIBLE" in r.log[-2]


def test_pump_time_limit_reports_not_found_or_found():
    m = read_mps(example_path("markshare2"))
    r = pump(m, time_limit=0.0)
    assert r.iterations == 0


def test_rounding_invariants_p0201():
    """test/runtests.jl:57-71."""
    m = read_mps(example_path("p0201"))
    r = pump(m, time_limit=60.0)
    assert r.found
    rng = np.random.Generator(np.random.PCG64(1))
    assert np.array_equal(Rounder(m).round_guarded(r.x, rng, 0), r.x)
    x = np.full(len(m.c), 0.5) + 0.01 * np.random.default_rng(2).random(len(m.c))
    a = Rounder(m, policy=None).round_guarded(x, np.random.Generator(np.random.PCG64(3)), 20)
    b = Rounder(m, policy=None).round_sequential(x, np.random.Generator(np.random.PCG64(3)), 20)
    assert np.array_equal(a, b)
    assert np.all(a == np.round(a))


def _tiny_models():
    rng = np.random.default_rng(5)
    out = []
    for k in range(12):
        n = int(rng.integers(3, 7))
        m = int(rng.integers(1, 4))
        A = rng.integers(-3, 4, size=(m, n)).astype(float)
        A[:, rng.integers(0, n)] += 1.0
        u = rng.integers(1, 3, size=n).astype(float)
        x0 = np.floor(rng.random(n) * (u + 1))
        eq = rng.random(m) < 0.3
        b = A @ x0 + np.where(eq, 0.0, rng.integers(0, 3, size=m))
        out.append(MILP(c=np.zeros(n), c0=0.0, A=sp.csc_matrix(A), b=b,
                        row_sense=np.where(eq, "eq", "le"), l=np.zeros(n), u=u,
                        int_idx=np.arange(n)))
    return out


@pytest.mark.parametrize("idx", range(12))
def test_propagation_is_sound_and_incremental(idx):
    """Brute force on tiny models: after fixing x_0 to each value and
    propagating, every feasible integer completion stays inside the domains,
    and the incremental row activities equal a fresh recomputation."""
    m = _tiny_models()[idx]
    n = len(m.c)
    d = PropagationData(m)
    feasible = [np.array(p, float) for p in itertools.product(*[range(int(u) + 1) for u in m.u])
                if is_feasible(m, np.array(p, float))]
    for v in range(int(m.u[0]) + 1):
        s = PropagationState(d, 10**9)
        s.fix(0, float(v))
        s.propagate()
        Lmin, Lmax = activities(d, s.l, s.u)
        np.testing.assert_allclose(s.L_min, Lmin, atol=1e-9)
        np.testing.assert_allclose(s.L_max, Lmax, atol=1e-9)
        for p in feasible:
            if p[0] == v:
                assert not s.infeasible
                assert all(s.l[j] - 1e-9 <= p[j] <= s.u[j] + 1e-9 for j in range(n))


def test_round_to_nearest_and_flip():
    m = G.toy2()
    assert list(round_to_nearest(np.array([2.5, 0.3]), m)) == [2.0, 0.3]   # ties to even
    assert list(round_to_nearest(np.array([7.2, 0.0]), m)) == [5.0, 0.0]   # kept in bounds
    xt = np.array([1.0, 0.0])
    flip_farthest(xt, np.array([1.9, 0.0]), m, 5)
    assert xt[0] == 2.0


def test_tight_projections_leave_the_fp32_form():
    """Projections asked for a relative KKT tolerance below what FP32
    iterates resolve run in the FP64 form (pk1 is mixed: its integral
    projections are tightened towards 1e-8, src/pump.jl:340-356)."""
    from types import SimpleNamespace

    from oracle import oracle as O
    from pdfp_amd.pump import FP32_TOLERANCE_FLOOR

    calls = []

    def make_state(*args, precision="fp32", **kw):
        s = O.OraclePDFPState(*args, **kw)
        s.precision = precision
        return s

    def solve(s, p, pw):
        calls.append((s.precision, min(p.eps_optimal_relative, p.eps_feasible_relative)))
        return O.pdfp_solve(s, p, pw)

    mod = SimpleNamespace(PDFPState=make_state, set_target=O.set_target,
                          warm_start=O.warm_start, pdfp_solve=solve,
                          original_primal_residual=O.original_primal_residual)
    m = read_mps(example_path("pk1"))
    r = feasibility_pump(m, projector=mod, root_lp=oracle_root_lp, time_limit=120.0,
                         precision="fp32", scaling="host", polish=False)
    assert r.found and is_feasible(m, r.x)
    assert calls and calls[0][0] == "fp32"
    assert all(prec == "fp64" for prec, eps in calls if eps < FP32_TOLERANCE_FLOOR)
    switched = [i for i, (prec, _) in enumerate(calls) if prec == "fp64"]
    if switched:
        assert all(prec == "fp64" for prec, _ in calls[switched[0]:])
        assert any("FP64 form" in line for line in r.log)


# ---- the vectorised host steps against the reference's element loops ----------

def _loop_round_to_nearest(x, milp):
    """src/pump.jl:53-62, element by element."""
    import math

    xt = np.array(x, dtype=np.float64)
    for j in milp.int_idx.tolist():
        v = float(np.rint(x[j]))
        if math.isfinite(milp.l[j]) and v < milp.l[j]:
            v = float(math.ceil(milp.l[j]))
        if math.isfinite(milp.u[j]) and v > milp.u[j]:
            v = float(math.floor(milp.u[j]))
        xt[j] = v
    return xt


def _loop_flip_farthest(xt, x, milp, TT):
    """src/pump.jl:121-135, element by element."""
    l, u = milp.l, milp.u
    distances = [(abs(x[j] - xt[j]), j) for j in milp.int_idx.tolist()]
    distances.sort(key=lambda t: t[0], reverse=True)
    for k in range(min(TT, len(distances))):
        d, j = distances[k]
        if d == 0.0:
            continue
        cur = xt[j]
        direction = 1.0 if x[j] >= cur else -1.0
        v = cur + direction
        if v < l[j] or v > u[j]:
            v = cur - direction
        if v < l[j] or v > u[j]:
            continue
        xt[j] = v


def _loop_perturb_binaries(xt, x, milp, rng, is_general):
    """The binary part of src/pump.jl:149-187, element by element."""
    l, u = milp.l, milp.u
    for j in milp.int_idx.tolist():
        if is_general[j]:
            continue
        rho = -0.3 + float(rng.random())
        if abs(x[j] - xt[j]) + max(rho, 0.0) > 0.5:
            cur = xt[j]
            direction = 1.0 if x[j] >= cur else -1.0
            v = cur + direction
            if v < l[j] or v > u[j]:
                v = cur - direction
            if v < l[j] or v > u[j]:
                continue
            xt[j] = v


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_vectorised_pump_steps_equal_element_loops(seed):
    from pdfp_amd.pump import RestartLadder, _differs, _distance, perturb

    rng = np.random.default_rng(seed)
    m = G.general_integer(60, 120, seed)
    m.u[::7] = np.inf                        # some unbounded integers
    m.l[::11] = -np.inf
    n = len(m.c)
    I = m.int_idx
    for _ in range(20):
        x = rng.normal(2.0, 3.0, n)
        x[::5] = np.round(x[::5]) + 0.5      # exact ties
        a, b = round_to_nearest(x, m), _loop_round_to_nearest(x, m)
        assert np.array_equal(a, b)
        xt = np.round(rng.normal(2.0, 3.0, n))
        d_loop = 0.0
        for j in I.tolist():
            d_loop += abs(x[j] - xt[j])
        assert _distance(x, xt, I) == d_loop
        assert _differs(a, xt, I.tolist()) == any(a[j] != xt[j] for j in I.tolist())
        assert not _differs(a, a.copy(), I)
        f1, f2 = xt.copy(), xt.copy()
        flip_farthest(f1, x, m, 15)
        _loop_flip_farthest(f2, x, m, 15)
        assert np.array_equal(f1, f2)
        # perturb with no general integers touches only the binary part
        mb = G.set_cover(40, 70, seed)
        xb = rng.random(70)
        tb = np.round(rng.random(70))
        isg = np.zeros(70, dtype=bool)
        p1, p2 = tb.copy(), tb.copy()
        r1, r2 = np.random.default_rng(seed), np.random.default_rng(seed)
        perturb(p1, xb, mb, r1, [], isg, RestartLadder(0.0, 0, 20, 10), 5)
        _loop_perturb_binaries(p2, xb, mb, r2, isg)
        assert np.array_equal(p1, p2) and r1.random() == r2.random()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_total_integer_row_violation_equals_row_loop(seed):
    """src/pump.jl:98-118: the CSC activity and the left-to-right sum over
    the rows without continuous variables, as a plain loop restates it."""
    from pdfp_amd.pump import total_integer_row_violation

    m = G.mixed(80, 160, seed) if seed == 1 else G.general_integer(80, 160, seed)
    rng = np.random.default_rng(seed)
    is_int = m.is_int
    A = m.A.tocsc()
    for _ in range(5):
        x = np.round(rng.normal(0.5, 2.0, A.shape[1]))
        act = np.zeros(A.shape[0])
        skip = np.zeros(A.shape[0], dtype=bool)
        for j in range(A.shape[1]):
            for k in range(A.indptr[j], A.indptr[j + 1]):
                i = A.indices[k]
                act[i] += A.data[k] * x[j]
                if not is_int[j]:
                    skip[i] = True
        total = 0.0
        for i in range(A.shape[0]):
            if skip[i]:
                continue
            d = act[i] - m.b[i]
            total += max(0.0, d) if m.row_sense[i] == "le" else abs(d)
        assert total_integer_row_violation(m, x) == total


@pytest.mark.parametrize("n_int,seed", [(10, 1), (11, 2), (977, 3), (70_000, 4)])
def test_ranking_swaps_follow_scalar_draws(n_int, seed):
    """The ranking's |I|/10 transpositions (src/rounding.jl:108-112), drawn in
    one vector call, equal the reference's loop of scalar draws
    ``a = rand(rng, 1:len); b = rand(rng, 1:len)`` applied one swap at a time."""
    from pdfp_amd.milp import MILP
    from pdfp_amd.rounding import Rounder

    rng0 = np.random.default_rng(seed)
    n = n_int + 3
    A = sp.random(4, n, density=0.5, random_state=seed, format="csc")
    milp = MILP(c=np.zeros(n), c0=0.0, A=A, b=np.ones(4), row_sense=np.array(["le"] * 4),
                l=np.zeros(n), u=np.full(n, 3.0), int_idx=np.arange(n_int))
    x = rng0.uniform(0.0, 3.0, n)
    r = Rounder(milp, backend="python")
    got = r.ranking(x, np.random.Generator(np.random.PCG64(seed)), 10)
    # the reference's loop, one scalar draw at a time
    ranked = r.ranking(x, None, 0)
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(max(1, n_int // 10)):
        a = int(rng.integers(1, n_int + 1)) - 1
        b = int(rng.integers(1, n_int + 1)) - 1
        ranked[a], ranked[b] = ranked[b], ranked[a]
    assert np.array_equal(got, ranked)
    assert sorted(got.tolist()) == list(range(n_int))
    assert not np.array_equal(got, r.ranking(x, None, 0))


# ---- the README's p0201 log (README.md:36-48) ------------------------------------

def _iteration_lines(r):
    return [line.split() for line in r.log if line[:6].strip().isdigit()]


def test_p0201_readme_log_through_iteration_one():
    """README.md:38-42 on the oracle projector and root LP: the root LP ends
    OPTIMAL after 600 iterations, the first rounding violates the rows by 9,
    and pump iteration 1 has ℓ1 distance 8.08613, violation 1 and 600 PDHG
    iterations, all exactly as printed."""
    m = read_mps(example_path("p0201"))
    r = pump(m, time_limit=60.0, batch_policy="always")
    assert "root LP     OPTIMAL after 600 iterations" in r.log[2]
    assert "(first rounding violates by 9)" in r.log[3]
    it1 = _iteration_lines(r)[0]
    assert it1[:5] == ["1", "8.08613", "1", "600", "1.00e-02"] and it1[5] == "round"
    assert r.found and is_feasible(m, r.x)


def _perturbed_module(start_rel, frac_rel, seed):
    """The oracle projector with the first projection's start point scaled by
    (1 + start_rel·U[-1,1]) and its result's fractional coordinates (those
    not exactly integral in original units) by (1 + frac_rel·U[-1,1])."""
    from types import SimpleNamespace

    from oracle import oracle as O
    from pdfp_amd.pdfp import projector_problem

    m = read_mps(example_path("p0201"))
    D_c = projector_problem(m, scaling="host").D_c
    rng = np.random.default_rng(seed)
    calls = {"n": 0}

    class State(O.OraclePDFPState):
        def set_x(self, x):
            x = np.asarray(x, dtype=np.float64)
            super().set_x(x * (1.0 + start_rel * rng.uniform(-1.0, 1.0, x.size)))

        @property
        def pdhg_x(self):
            v = self.vector("pdhg_x")
            calls["n"] += 1
            if calls["n"] == 1:
                x = D_c * v
                frac = np.abs(x - np.rint(x)) != 0.0
                v = np.where(frac, v * (1.0 + frac_rel * rng.uniform(-1.0, 1.0, v.size)), v)
            return v

    return SimpleNamespace(PDFPState=State, set_target=O.set_target, warm_start=O.warm_start,
                           pdfp_solve=O.pdfp_solve,
                           original_primal_residual=O.original_primal_residual)


@pytest.mark.parametrize("start_rel,frac_rel", [(1e-6, 0.0), (0.0, 1e-4), (1e-9, 1e-3)])
def test_p0201_second_target_is_not_an_arithmetic_knife_edge(start_rel, frac_rel):
    """The README's iteration 2 (1.23237 / 5 / 400) differs from the
    restatement's (1.15587 / 1 / 600).  The second target is the rounding of
    projection 1, so this pins what does NOT explain the difference:
    perturbations far larger than summation-order or FMA effects — the
    projection's start point by up to 1e-6 relative, the fractional
    coordinates of projection 1 (whose ranking keys lie 4e-6 apart) by up to
    1e-3 — leave iteration 1 and iteration 2 unchanged."""
    m = read_mps(example_path("p0201"))
    for seed in range(3):
        r = feasibility_pump(m, projector=_perturbed_module(start_rel, frac_rel, seed),
                             root_lp=oracle_root_lp, scaling="host", batch_policy="always",
                             time_limit=60.0)
        lines = _iteration_lines(r)
        # the printed distance of iteration 1 is measured on the perturbed point
        assert lines[0][2:4] == ["1", "600"]
        assert lines[0][1] == "8.08613" or frac_rel > 0
        assert lines[1][:4] == ["2", "1.15587", "1", "600"]


def test_fp64_polish_finishes_each_fp32_projection():
    """feasibility_pump(polish=True) on the FP32 form: every FP32 projection
    is followed by an FP64 solve started at the FP32 point (x, y and z), with
    the same target and tolerances, and the rounder receives the FP64 point
    (the oracle stands in for both forms here)."""
    from types import SimpleNamespace

    from oracle import oracle as O

    events = []

    class State(O.OraclePDFPState):
        def __init__(self, *args, precision="fp32", **kw):
            super().__init__(*args, **kw)
            self.precision = precision

        def set_dual(self, y, z):
            events.append(("set_dual", self.precision))
            super().set_dual(y, z)

    def solve(s, p, pw):
        events.append(("solve", s.precision, p.eps_optimal_relative, p.eps_feasible_relative))
        return O.pdfp_solve(s, p, pw)

    mod = SimpleNamespace(PDFPState=State, set_target=O.set_target, warm_start=O.warm_start,
                          pdfp_solve=solve, original_primal_residual=O.original_primal_residual)
    m = G.mixed(300, 600, 2)
    r = feasibility_pump(m, projector=mod, root_lp=oracle_root_lp, time_limit=120.0,
                         precision="fp32", scaling="host", polish=True, batch_policy="always")
    assert r.found and is_feasible(m, r.x)
    solves = [e for e in events if e[0] == "solve"]
    fp32 = [i for i, e in enumerate(solves) if e[1] == "fp32"]
    assert fp32
    for i in fp32:      # each FP32 solve is finished by an FP64 one, same tolerances
        assert solves[i + 1][1] == "fp64" and solves[i + 1][2:] == solves[i][2:]
    assert events.count(("set_dual", "fp64")) == len(fp32)
