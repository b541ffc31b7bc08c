"""The device preconditioner (pdfp_scale_matrix) against the oracle's
scale_problem (src/preconditioner.jl:26-99), bit for bit.

Bar: integer-exact comparison of every output (D_r, D_c, the scaled values
and the scaled vectors) — the arithmetic is max, sqrt, divide and multiply
in the reference's order, and each Pock–Chambolle sum is accumulated in the
reference's order, so nothing is allowed to differ.  At the 1M-nonzero size
the host numpy restatement (itself pinned bitwise to the oracle on the CPU
suite) is the checker.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd.milp import read_mps
from pdfp_amd.preconditioner import scale_problem

from .conftest import example_path

pytestmark = pytest.mark.gpu

KEYS = ("D_r", "D_c", "c", "var_lb", "var_ub", "con_lb", "con_ub")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _with_empty_row_and_column():
    m = G.set_cover(60, 90, 5)
    A = m.A.tolil()
    A[7, :] = 0          # an empty row
    A[:, 11] = 0         # an empty column
    m.A = sp.csc_matrix(A)
    m.A.eliminate_zeros()
    return m


MAKERS = {
    "set_cover": lambda: G.set_cover(133, 201, 2),
    "mixed": lambda: G.mixed(500, 900, 4),
    "power_law": lambda: G.mixed(2000, 5000, 3, power_law=True, nnz=60_000, max_len=3000),
    "general_integer": lambda: G.general_integer(800, 1600, 2),
    "p0201": lambda: read_mps(example_path("p0201")),
    "pk1": lambda: read_mps(example_path("pk1")),
    "empty_row_col": _with_empty_row_and_column,
}


@pytest.mark.parametrize("rescale", [False, True])
@pytest.mark.parametrize("name", list(MAKERS))
def test_device_scaling_equals_oracle_bitwise(name, rescale):
    say("scale", name, rescale)
    m = MAKERS[name]()
    con_lb = np.where(m.row_sense == "eq", m.b, -np.inf)
    a = scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, bound_objective_rescaling=rescale,
                      backend="device")
    b = O.scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, bound_objective_rescaling=rescale)
    for k in KEYS:
        assert np.array_equal(getattr(a, k), b[k]), k
    assert np.array_equal(a.A.data, b["A"].data)
    assert np.array_equal(a.A.indices, b["A"].indices)
    assert a.bound_scale == b["bound_scale"] and a.objective_scale == b["objective_scale"]


@pytest.mark.parametrize("ruiz", [0, 1, 3])
def test_device_scaling_ruiz_rounds(ruiz):
    say("scale ruiz", ruiz)
    m = G.mixed(300, 700, 9)
    con_lb = np.full(300, -np.inf)
    a = scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, ruiz_iterations=ruiz, backend="device")
    b = O.scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, ruiz_iterations=ruiz)
    assert np.array_equal(a.A.data, b["A"].data)
    assert np.array_equal(a.D_r, b["D_r"]) and np.array_equal(a.D_c, b["D_c"])


def test_device_scaling_rejects_other_alpha():
    from pdfp_amd._lib import LibraryError

    m = G.set_cover(40, 60, 1)
    with pytest.raises(LibraryError, match="alpha"):
        scale_problem(m.A, m.c, m.l, m.u, np.full(40, -np.inf), m.b, pock_chambolle_alpha=0.5,
                      backend="device")


def test_device_scaling_equals_host_at_1m_nonzeros():
    say("scale 1M")
    m = G.mixed(50_000, 100_000, 1)
    con_lb = np.full(50_000, -np.inf)
    a = scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, backend="device")
    h = scale_problem(m.A, m.c, m.l, m.u, con_lb, m.b, backend="host")
    assert np.array_equal(a.A.data, h.A.data)
    for k in KEYS:
        assert np.array_equal(getattr(a, k), getattr(h, k)), k
