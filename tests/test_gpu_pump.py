"""The pump (src/pump.jl) with the device root LP and device projector.

- test/runtests.jl:73-79 on the device: the pump finds a feasible point on
  the four example instances, in both projector forms.
- README.md:36-42 (p0201): the root LP ends OPTIMAL after 600 iterations,
  the first rounding violates the rows by 9, and pump iteration 1 has ℓ1
  distance 8.08613, violation 1 and 600 PDHG iterations — the reference's
  log, reproduced by the FP64 form (iteration 2 differs; see the oracle's
  tests of the same log).
- The FP64 polish of the FP32 form keeps the pump's end point feasible.
"""
import re

import numpy as np
import pytest

from pdfp_amd import generators as G
from pdfp_amd.milp import is_feasible, read_mps
from pdfp_amd.pump import feasibility_pump

from .conftest import example_path

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["p0201", "gen-ip054", "pk1", "markshare2"])
def test_pump_on_examples_device(name, precision):
    print("pump", name, precision, flush=True)
    m = read_mps(example_path(name))
    r = feasibility_pump(m, time_limit=120.0, precision=precision)
    assert r.found
    assert is_feasible(m, r.x)
    assert abs(r.objective - (float(m.c @ r.x) + m.c0)) <= 1e-9 * (1 + abs(r.objective))


def test_p0201_log_matches_reference_readme():
    m = read_mps(example_path("p0201"))
    r = feasibility_pump(m, time_limit=60.0, precision="fp64")
    log = "".join(r.log)
    assert re.search(r"root LP\s+OPTIMAL after 600 iterations", log)
    assert "first rounding violates by 9)" in log
    it1 = [line for line in r.log if re.match(r"\s+1\s", line)][0].split()
    assert it1[1] == "8.08613" and it1[2] == "1" and it1[3] == "600"


@pytest.mark.parametrize("name", ["p0201", "gen-ip054", "pk1", "markshare2"])
def test_batched_pump_finds_feasible(name):
    """K = 8 trajectories (seeds 100k + 1) on one batched projector, the
    first feasible point wins (replaces scripts/solve.jl:121-148)."""
    from pdfp_amd.pump import feasibility_pump_batch

    print("batched pump", name, flush=True)
    m = read_mps(example_path(name))
    r = feasibility_pump_batch(m, trajectories=8, time_limit=120.0)
    assert r.found and is_feasible(m, r.x)
    assert 0 <= r.trajectory < 8 and r.seed == 100 * (r.trajectory + 1) + 1
    assert abs(r.objective - (float(m.c @ r.x) + m.c0)) <= 1e-9 * (1 + abs(r.objective))


def test_batched_pump_mixed_model_takes_fp64_projections():
    """A mixed model whose integral projections are tightened below the FP32
    form's reach moves those trajectories to FP64 projectors."""
    from pdfp_amd import generators as G
    from pdfp_amd.pump import feasibility_pump_batch

    print("batched pump mixed", flush=True)
    m = G.mixed(2000, 4000, 1)
    r = feasibility_pump_batch(m, trajectories=4, time_limit=120.0)
    assert r.found and is_feasible(m, r.x)


def test_pack_pump_finds_feasible_for_every_instance():
    """The four example instances and a mixed model, each with its own pump,
    all projections in one block-diagonal pack (PDFPPack)."""
    from pdfp_amd import generators as G
    from pdfp_amd.pump import feasibility_pump_pack

    print("pack pump", flush=True)
    milps = [read_mps(example_path(n)) for n in ("p0201", "gen-ip054", "pk1", "markshare2")]
    milps.append(G.mixed(2000, 4000, 2))
    rs = feasibility_pump_pack(milps, time_limit=120.0)
    assert len(rs) == len(milps)
    for m, r in zip(milps, rs):
        assert r.found and is_feasible(m, r.x), m.name
        assert abs(r.objective - (float(m.c @ r.x) + m.c0)) <= 1e-9 * (1 + abs(r.objective))


@pytest.mark.parametrize("seed", [1, 2])
def test_pump_fp32_with_fp64_polish(seed):
    """feasibility_pump(precision="fp32", polish=True): each FP32 projection
    is finished by an FP64 solve from its point; the pump still ends with a
    point that is_feasible accepts."""
    print("pump polish", seed, flush=True)
    m = G.mixed(2000, 4000, 1)
    r = feasibility_pump(m, seed=seed, time_limit=60.0, precision="fp32", polish=True)
    assert r.found and is_feasible(m, r.x)
