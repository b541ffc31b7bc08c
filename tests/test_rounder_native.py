"""The C++ rounder (pdfp_rounder_*, host code in libpdfp_hip) against the
Python restatement of src/rounding.jl and src/propagation.jl, bit for bit.

Both backends receive the same LP point and draw the same random numbers
(the ranking stays in Python), so every output must be identical: CPR,
the guarded two-phase CPR (both schedules, the guard's fallback), the
violation-guided rule and the propagation budget.  The Python restatement is
itself checked against the reference's p0201 log in test_pump_cpu.py.
"""
import numpy as np
import pytest

from pdfp_amd import generators as G
from pdfp_amd.milp import read_mps
from pdfp_amd.rounding import BatchPolicy, Rounder

from .conftest import example_path


def _instances():
    yield "p0201", read_mps(example_path("p0201"))
    yield "pk1", read_mps(example_path("pk1"))
    yield "gen-ip054", read_mps(example_path("gen-ip054"))
    yield "markshare2", read_mps(example_path("markshare2"))
    yield "set_cover", G.set_cover(300, 500, 4)
    yield "mixed", G.mixed(400, 900, 6)
    yield "general_integer", G.general_integer(300, 600, 3)


def _points(milp, seed, k=4):
    """LP-like points: integral on some coordinates, fractional on others."""
    rng = np.random.default_rng(seed)
    n = len(milp.c)
    lo = np.where(np.isfinite(milp.l), milp.l, -5.0)
    hi = np.where(np.isfinite(milp.u), milp.u, lo + 10.0)
    for _ in range(k):
        x = lo + rng.random(n) * (hi - lo)
        snap = rng.random(n) < 0.6
        x[snap] = np.round(x[snap])
        yield x


@pytest.mark.parametrize("name,milp", list(_instances()), ids=lambda v: v if isinstance(v, str) else "")
@pytest.mark.parametrize("schedule", ["all", "regroup"])
@pytest.mark.parametrize("guided", [False, True])
def test_native_rounder_matches_python(name, milp, schedule, guided):
    kw = dict(schedule=schedule, policy=None)
    native = Rounder(milp, backend="native", **kw)
    python = Rounder(milp, backend="python", **kw)
    if guided:
        d = [False] * len(milp.c)
        for j in milp.int_idx.tolist():
            d[j] = True
        native.direction = d
        python.direction = list(d)
    for t, x in enumerate(_points(milp, 11)):
        for it in (0, 12):                      # before and after swap_iteration
            r1, r2 = np.random.default_rng(t), np.random.default_rng(t)
            a = native.round_guarded(x, r1, it)
            b = python.round_guarded(x, r2, it)
            assert np.array_equal(a, b), (name, t, it)
            assert r1.random() == r2.random()   # the same draws were made
            a = native.round_sequential(x, np.random.default_rng(t), it)
            b = python.round_sequential(x, np.random.default_rng(t), it)
            assert np.array_equal(a, b), (name, t, it)
    assert (native.n_calls, native.n_batched, native.n_fallbacks) == \
        (python.n_calls, python.n_batched, python.n_fallbacks)


def test_native_rounder_budget_and_fallback():
    """A small propagation budget (max_pops) and a batch that the guard
    discards take the same paths in both backends."""
    milp = read_mps(example_path("p0201"))
    native = Rounder(milp, backend="native", policy=None, max_pops=5)
    python = Rounder(milp, backend="python", policy=None, max_pops=5)
    for t, x in enumerate(_points(milp, 3, k=8)):
        a = native.round_guarded(x, np.random.default_rng(t), 0)
        b = python.round_guarded(x, np.random.default_rng(t), 0)
        assert np.array_equal(a, b)
    assert native.n_fallbacks == python.n_fallbacks


def test_native_rounder_rejects_bad_ranking():
    from pdfp_amd._lib import LibraryError

    milp = G.set_cover(20, 30, 1)
    r = Rounder(milp, backend="native", policy=None)
    with pytest.raises(LibraryError, match="ranked"):
        r._native.round_sequential(np.zeros(30), np.array([0, 31]))


def test_native_rounder_with_policy_runs():
    milp = read_mps(example_path("pk1"))
    r = Rounder(milp, backend="native", policy=BatchPolicy())
    x = next(_points(milp, 5))
    for it in range(30):
        xt = r.round_guarded(x, np.random.default_rng(it), it)
        I = milp.int_idx
        assert np.array_equal(xt[I], np.round(xt[I]))
    assert r.n_calls == 30
