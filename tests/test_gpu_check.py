"""The device feasibility check (pdfp_checker_*) against max_violation
(src/milp.jl:102-114), bit for bit: each row's A x is summed in the order of
the reference's CSC product on both sides, so the values must be equal, not
close."""
import numpy as np
import pytest

from pdfp_amd import generators as G
from pdfp_amd.milp import DeviceViolationCheck, max_violation, read_mps

from .conftest import example_path

pytestmark = pytest.mark.gpu


def _points(milp, seed, k=6):
    rng = np.random.default_rng(seed)
    n = len(milp.c)
    lo = np.where(np.isfinite(milp.l), milp.l, -5.0)
    hi = np.where(np.isfinite(milp.u), milp.u, lo + 10.0)
    for t in range(k):
        x = lo + rng.random(n) * (hi - lo)
        if t % 2:
            x = np.round(x)
        yield x


@pytest.mark.parametrize("name,make", [
    ("p0201", lambda: read_mps(example_path("p0201"))),
    ("pk1", lambda: read_mps(example_path("pk1"))),
    ("markshare2", lambda: read_mps(example_path("markshare2"))),
    ("mixed", lambda: G.mixed(3000, 6000, 2)),
    ("general_integer", lambda: G.general_integer(2000, 4000, 5)),
    ("power_law", lambda: G.mixed(2000, 5000, 3, power_law=True, nnz=60_000, max_len=3000)),
])
def test_device_check_equals_host_bitwise(name, make):
    print("check", name, flush=True)
    m = make()
    chk = DeviceViolationCheck(m)
    for x in _points(m, 7):
        assert chk(x) == max_violation(m, x)
    # a feasible point: the generators' x0 is not kept, so use the zero
    # vector on a set cover, which violates every row by exactly 1
    sc = G.set_cover(50, 80, 3)
    assert DeviceViolationCheck(sc)(np.zeros(80)) == max_violation(sc, np.zeros(80)) == 1.0


def test_device_check_rejects_wrong_length():
    m = G.set_cover(20, 30, 1)
    with pytest.raises(ValueError):
        DeviceViolationCheck(m)(np.zeros(29))
