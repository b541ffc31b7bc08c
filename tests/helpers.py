"""Shared test helpers: fixture instances and a projector driver that works
with either the HIP state or the oracle state."""
from __future__ import annotations

import json
import os

import numpy as np

from pdfp_amd import generators as G
from pdfp_amd.milp import read_mps
from pdfp_amd.pdfp import SolverParameters, projector_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def build_instance(spec: dict):
    f = spec["family"]
    if f == "toy1":
        return G.toy1()
    if f == "toy2":
        return G.toy2()
    if f == "set_cover":
        return G.set_cover(spec["m"], spec["n"], spec["seed"])
    if f == "mixed":
        return G.mixed(spec["m"], spec["n"], spec["seed"])
    if f == "general_integer":
        return G.general_integer(spec["m"], spec["n"], spec["seed"])
    if f == "example":
        return read_mps(os.path.join(GOLDEN, "examples", spec["name"] + ".mps.gz"))
    raise KeyError(f)


def l1_fixtures():
    with open(os.path.join(GOLDEN, "l1_projections.json")) as f:
        return json.load(f)


def root_lp_fixtures():
    with open(os.path.join(GOLDEN, "root_lp.json")) as f:
        return json.load(f)


def project_with(module, state_cls, milp, x_tilde, tol=1e-8, iteration_limit=1_000_000,
                 **kw):
    """The reference's test helper (test/runtests.jl:12-27) with a chosen
    state implementation: module provides set_target / pdfp_solve."""
    # the oracle's runs stay on the CPU: its scaling is the host restatement
    on_host = getattr(module, "__name__", "").startswith("oracle")
    pp = projector_problem(milp, scaling="host" if on_host else "device")
    s = state_cls(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w, **kw)
    module.set_target(s, np.asarray(x_tilde, dtype=np.float64) / pp.D_c)
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=iteration_limit)
    status = module.pdfp_solve(s, p, 1.0)
    return pp, s, status


def l1_distance(milp, x, x_tilde) -> float:
    I = milp.int_idx
    return float(np.abs(np.asarray(x)[I] - np.asarray(x_tilde)[I]).sum())


def oracle_projector_module():
    """The oracle with the projector-module interface the pump expects (tests
    only: it stands in for the device on the CPU)."""
    from types import SimpleNamespace

    from oracle import oracle as O

    return SimpleNamespace(PDFPState=O.OraclePDFPState, set_target=O.set_target,
                           warm_start=O.warm_start, pdfp_solve=O.pdfp_solve,
                           original_primal_residual=O.original_primal_residual)


def oracle_root_lp(milp, *, tolerance, time_limit, cupdlpx_termination):
    from oracle import oracle as O

    return O.solve_lp_relaxation(milp, tolerance=tolerance, time_limit=time_limit,
                                 cupdlpx_termination=cupdlpx_termination)
