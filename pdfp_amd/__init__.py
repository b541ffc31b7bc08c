"""pdfp_amd — an MI355X-native PDHG projector and root-LP solver for the PDFP
feasibility pump.

The device work (CSR passes with fused PDHG epilogues, FP64 reductions) lives
in libpdfp_hip.so (C ABI: include/pdfp_hip.h); this package is the host-side
mirror of the reference's Julia API (MIT-Lu-Lab/PDFP, src/PDFP.jl:22-23).
"""
from .milp import (MILP, is_feasible, is_integral, max_violation, num_constraints,
                   num_continuous, num_variables, read_mps, write_solution)

__all__ = ["MILP", "read_mps", "max_violation", "is_feasible", "write_solution",
           "is_integral", "num_variables", "num_constraints", "num_continuous"]


def __getattr__(name):
    # The solver pieces load libpdfp_hip on first use (and raise if it is missing).
    if name in ("feasibility_pump", "PumpResult"):
        from . import pump
        return getattr(pump, name)
    raise AttributeError(name)
