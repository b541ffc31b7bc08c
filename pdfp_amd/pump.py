"""The feasibility pump with the device projector (host loop).

Restates src/pump.jl: ``feasibility_pump(milp; ...) -> PumpResult`` with the
same keyword arguments and defaults.  The root LP runs on the device
(pdfp_amd.pdlp), every ℓ1 projection runs on the device (pdfp_amd.pdfp,
scaled data, target and iterates resident in HBM, warm-started), and the
rounding, propagation, flips, perturbations and cycle detection stay on the
CPU with the reference's arithmetic.

Differences that cannot be avoided without Julia:
- random numbers come from numpy's PCG64 seeded with ``seed`` instead of
  Julia's MersenneTwister (the same draws at the same points);
- cycle detection keeps the exact integer coordinates of each point in a set
  instead of Julia's 64-bit ``hash`` (identical except on hash collisions).
"""
from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from .milp import MILP, is_integral, max_violation, num_constraints, num_continuous, num_variables
from .rounding import BatchPolicy, Rounder, jround

# Smallest relative KKT tolerance handed to the FP32 projector form.
FP32_TOLERANCE_FLOOR = 1e-5


@dataclass
class PumpResult:
    """Outcome of feasibility_pump (src/pump.jl:22-35).
    time-to-first-feasible is ``time`` when ``found``."""

    found: bool
    x: np.ndarray
    objective: float
    max_violation: float
    iterations: int
    perturbations: int
    pdhg_iterations: int
    root_lp_status: str
    root_lp_time: float
    projection_time: float
    rounding_time: float
    time: float
    log: list = field(default_factory=list, repr=False)


@dataclass
class _Counters:
    iterations: int = 0
    perturbations: int = 0
    pdhg_iterations: int = 0
    root_lp_status: str = "UNSPECIFIED"
    root_lp_time: float = 0.0
    projection_time: float = 0.0
    rounding_time: float = 0.0


def _differs(a, b, idx) -> bool:
    return any(a[j] != b[j] for j in idx)


def round_to_nearest(x, milp: MILP) -> np.ndarray:
    """Nearest integer on I, kept inside the bounds (src/pump.jl:53-62)."""
    xt = np.array(x, dtype=np.float64)
    for j in milp.int_idx.tolist():
        v = jround(x[j])
        if math.isfinite(milp.l[j]) and v < milp.l[j]:
            v = float(math.ceil(milp.l[j]))
        if math.isfinite(milp.u[j]) and v > milp.u[j]:
            v = float(math.floor(milp.u[j]))
        xt[j] = v
    return xt


def integer_key(x, int_idx) -> bytes:
    """Exact identity of the integer coordinates (the reference hashes them,
    src/pump.jl:66-74; -0.0 is mapped to 0.0 first)."""
    v = np.asarray(x, dtype=np.float64)[int_idx] + 0.0
    return v.tobytes()


def batch_schedule(milp: MILP, x, dual_slack, push_threshold: float) -> str:
    """src/pump.jl:81-93."""
    n_pushed = 0
    n_integral = 0
    for j in milp.int_idx.tolist():
        if not abs(x[j] - jround(x[j])) == 0.0:
            continue
        n_integral += 1
        at_lower = math.isfinite(milp.l[j]) and abs(x[j] - milp.l[j]) < 1e-9
        at_upper = math.isfinite(milp.u[j]) and abs(x[j] - milp.u[j]) < 1e-9
        if (at_lower and dual_slack[j] > 0) or (at_upper and dual_slack[j] < 0):
            n_pushed += 1
    fraction = n_pushed / n_integral if n_integral > 0 else 1.0
    return "regroup" if fraction < push_threshold else "all"


def total_integer_row_violation(milp: MILP, x) -> float:
    """Total violation of the rows without continuous variables
    (src/pump.jl:98-118)."""
    A = milp.A
    m = num_constraints(milp)
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    activity = np.bincount(A.indices, weights=A.data * np.asarray(x)[cols], minlength=m)
    is_int = milp.is_int
    skip = np.zeros(m, dtype=bool)
    cont_entries = ~is_int[cols]
    skip[A.indices[cont_entries]] = True
    total = 0.0
    b = milp.b
    le = milp.row_sense == "le"
    for i in range(m):
        if skip[i]:
            continue
        total += max(0.0, activity[i] - b[i]) if le[i] else abs(activity[i] - b[i])
    return total


def flip_farthest(xt: np.ndarray, x, milp: MILP, TT: int) -> None:
    """FP's stalling move: flip the TT integer variables farthest from x
    (src/pump.jl:121-135)."""
    l, u = milp.l, milp.u
    distances = [(abs(x[j] - xt[j]), j) for j in milp.int_idx.tolist()]
    distances.sort(key=lambda t: t[0], reverse=True)           # stable, as sort!(rev=true)
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


@dataclass
class RestartLadder:
    """src/pump.jl:139-144."""

    flips: float
    last_restart: int
    average_flips: int
    max_flips: int


def perturb(xt: np.ndarray, x, milp: MILP, rng, general: list, is_general: np.ndarray,
            ladder: RestartLadder, iteration: int) -> None:
    """FP's restart (src/pump.jl:149-187)."""
    l, u = milp.l, milp.u
    if general:
        while ladder.last_restart < iteration:
            ladder.last_restart += 1
            ladder.flips *= 0.85
        ladder.flips = min(ladder.flips + 2 * ladder.average_flips + 1, ladder.max_flips)
        for _ in range(max(1, int(math.floor(ladder.flips)))):
            j = general[int(rng.integers(1, len(general) + 1)) - 1]
            lo, hi = l[j], u[j]
            r = float(rng.random())
            if math.isfinite(hi - lo) and hi - lo < 1e15:
                v = math.floor(lo + (1.0 + hi - lo) * r)
            elif math.isfinite(lo) and xt[j] - lo < 1e9:
                v = lo + (2e9 - 1.0) * r
            elif math.isfinite(hi) and hi - xt[j] < 1e9:
                v = hi - (2e9 - 1.0) * r
            else:
                v = xt[j] + (2e9 - 1.0) * r - 1e9
            v = jround(hi if v > hi else (lo if v < lo else v))
            if abs(v - xt[j]) > 1e-6:
                xt[j] = v
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


def feasibility_pump(milp: MILP, *, time_limit: float = 1800.0, seed: int = 1,
                     verbose: bool = False, flip_parameter: int = 20,
                     restart_period: int = 100, swap_iteration: int = 10,
                     gap_start: float = 1e-2, gap_decay: float = 0.98, gap_floor: float = 1e-4,
                     residual_ratio: float = 0.1, root_lp_tolerance: float = 1e-2,
                     projection_iteration_limit: int = 500_000, batch_policy: str = "cost",
                     push_threshold: float = 0.01, violation_threshold: float = 5000.0,
                     propagation_budget: int = 100_000, feasibility_tolerance: float = 1e-6,
                     precision: str = "fp32", device: int = 0, scaling: str = "device",
                     projector=None, root_lp=None) -> PumpResult:
    """Search for a feasible solution of ``milp`` with the feasibility pump
    (src/pump.jl:231-410).  The keyword arguments and defaults are the
    reference's (README.md:89-108); ``precision`` picks the projector's form
    ("fp32": FP32 data and iterates with FP64 reductions and polish; "fp64");
    ``scaling`` where the projector's matrix is equilibrated ("device" or
    "host").

    ``projector`` / ``root_lp`` replace the device solvers: a module with
    ``PDFPState``, ``set_target``, ``warm_start``, ``pdfp_solve``,
    ``original_primal_residual`` and a ``solve_lp_relaxation`` callable.  They
    default to the device (pdfp_amd.pdfp, pdfp_amd.pdlp); there is no CPU
    fallback.
    """
    if batch_policy not in ("cost", "always"):
        raise ValueError("batch_policy must be 'cost' or 'always'")
    from . import pdfp as _pdfp_mod
    from . import pdlp as _pdlp_mod

    P = projector if projector is not None else _pdfp_mod
    if root_lp is not None:
        solve_lp = root_lp
    else:
        def solve_lp(milp_, **kw):
            return _pdlp_mod.solve_lp_relaxation(milp_, scaling=scaling, device=device, **kw)
    t0 = time.perf_counter()

    def elapsed():
        return time.perf_counter() - t0

    def remaining():
        return max(time_limit - elapsed(), 1.0)

    n = num_variables(milp)
    m = num_constraints(milp)
    int_idx = milp.int_idx
    int_list = int_idx.tolist()
    tol = feasibility_tolerance
    mixed = num_continuous(milp) > 0
    rng = np.random.Generator(np.random.PCG64(seed))
    rounder = Rounder(milp, max_pops=(2**63 - 1) if mixed else propagation_budget,
                      swap_iteration=swap_iteration,
                      policy=BatchPolicy() if batch_policy == "cost" else None)
    stats = _Counters()
    log: list = []

    def emit(s: str):
        log.append(s)
        if verbose:
            sys.stdout.write(s)
            sys.stdout.flush()

    emit(f"PDFP feasibility pump: {milp.name}\n")
    emit(f"model       {m} rows, {n} columns ({len(int_list)} integer, {num_continuous(milp)} "
         f"continuous), {milp.A.nnz} nonzeros\n")

    # Root LP relaxation.
    x, dual_slack, stats.root_lp_status, root_iterations = solve_lp(
        milp, tolerance=root_lp_tolerance, time_limit=remaining(), cupdlpx_termination=mixed)
    stats.root_lp_time = elapsed()
    emit(f"root LP     {stats.root_lp_status} after {root_iterations} iterations, "
         f"{stats.root_lp_time:.2f} s\n")
    if is_integral(x, int_idx, tol):
        xt = round_to_nearest(x, milp)
        if max_violation(milp, xt) <= tol:
            return _finish(milp, True, xt, stats, elapsed(), emit, log)

    # First rounding, then the decision on the violation-guided rule.
    rounder.schedule = batch_schedule(milp, x, dual_slack, push_threshold)
    t = time.perf_counter()
    xt = rounder.round_guarded(x, rng, 0)
    stats.rounding_time += time.perf_counter() - t
    total_violation = total_integer_row_violation(milp, xt)
    if total_violation >= violation_threshold:
        direction = [False] * n
        for j in int_list:
            direction[j] = True
        rounder.direction = direction
    emit(f"rounder     {rounder.schedule} schedule, violation-guided rounding "
         f"{'off' if rounder.direction is None else 'on'} (first rounding violates by "
         f"{total_violation:.4g})\n")

    # The projector, built once: scaled data on the device, warm-started at
    # the root LP solution (src/pump.jl:293-309).
    pp = _pdfp_mod.projector_problem(milp, scaling=scaling, device=device)
    D_c = pp.D_c
    projector_state = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                                  precision=precision, device=device)
    if mixed:
        P.original_primal_residual(projector_state, pp.D_r)
    projector_state.set_x(x / D_c)
    primal_weight = math.sqrt(1.0 + (len(int_list) / m if m > 0 else 0.0))
    gap = gap_start
    tightening_budget = gap_start

    general = [j for j in int_list if not (milp.l[j] == 0.0 and milp.u[j] == 1.0)]
    is_general = np.zeros(n, dtype=bool)
    is_general[general] = True
    ladder = RestartLadder(0.0, 0, flip_parameter, max(len(general) // 10, 10))
    seen: set = set()
    h_current = integer_key(xt, int_idx)
    since_restart = 0

    emit(f"{'iter':>6s} {'distance':>12s} {'violation':>10s} {'pdhg':>10s} {'gap tol':>9s} "
         f"{'action':>8s} {'time':>9s}\n")
    params_cls = _pdfp_mod.SolverParameters
    proj = {"state": projector_state}

    def solve_projection(params, target_scaled=None):
        # FP32 iterates do not resolve relative KKT errors much below 1e-5:
        # from the first projection asked for less, the FP64 form takes over,
        # started at the last FP32 projection.
        st = proj["state"]
        if (getattr(st, "precision", None) == "fp32" and
                min(params.eps_optimal_relative,
                    params.eps_feasible_relative) < FP32_TOLERANCE_FLOOR):
            st64 = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                               precision="fp64", device=device)
            if mixed:
                P.original_primal_residual(st64, pp.D_r)
            st64.set_x(st.pdhg_x)
            P.set_target(st64, proj["target"] if target_scaled is None else target_scaled)
            emit(f"projector   FP64 form from gap tolerance "
                 f"{params.eps_optimal_relative:.2e} on\n")
            proj["state"] = st = st64
        elif target_scaled is not None:
            P.set_target(st, target_scaled)
        else:
            P.warm_start(st)
        if target_scaled is not None:
            proj["target"] = target_scaled
        P.pdfp_solve(st, params, primal_weight)
        return D_c * st.pdhg_x, st.total_count

    while elapsed() < time_limit:
        violation = max_violation(milp, xt)
        if violation <= tol and is_integral(xt, int_idx, tol):
            return _finish(milp, True, xt.copy(), stats, elapsed(), emit, log)
        stats.iterations += 1
        iteration = stats.iterations
        since_restart += 1

        # Project x̃ onto the LP relaxation.
        t = time.perf_counter()
        gap_used = gap
        params = params_cls(iteration_limit=projection_iteration_limit,
                            eps_optimal_relative=gap, eps_feasible_relative=gap * residual_ratio,
                            time_limit=remaining())
        x, inner = solve_projection(params, xt / D_c)
        # Mixed models: re-solve an integral projection whose rounding is
        # infeasible only through its continuous part (src/pump.jl:340-356).
        while (mixed and tightening_budget > 1e-6 and elapsed() < time_limit
               and is_integral(x, int_idx, tol)
               and max_violation(milp, round_to_nearest(x, milp)) > tol):
            tightening_budget *= 0.1
            gap = max(gap * 0.1, 1e-8)
            params = params_cls(iteration_limit=projection_iteration_limit,
                                eps_optimal_relative=gap,
                                eps_feasible_relative=gap * residual_ratio,
                                time_limit=remaining())
            x, count = solve_projection(params)
            inner += count
        stats.pdhg_iterations += inner
        if gap > gap_floor:
            gap = max(gap * gap_decay, gap_floor)
        stats.projection_time += time.perf_counter() - t
        distance = 0.0
        for j in int_list:
            distance += abs(x[j] - xt[j])

        if is_integral(x, int_idx, tol):
            candidate = round_to_nearest(x, milp)
            if max_violation(milp, candidate) <= tol:
                _print_iteration(emit, iteration, distance, violation, inner, gap_used,
                                 "integral", elapsed())
                return _finish(milp, True, candidate, stats, elapsed(), emit, log)

        # Round x; if the rounding did not change, flip (FP's stalling move).
        t = time.perf_counter()
        x_round = rounder.round_guarded(x, rng, iteration)
        stats.rounding_time += time.perf_counter() - t
        candidate = xt.copy()
        action = "round"
        if _differs(x_round, xt, int_list):
            candidate[:] = x_round
        else:
            flips = int(rng.integers(flip_parameter // 2, 3 * flip_parameter // 2 + 1))
            flip_farthest(candidate, x, milp, flips)
            action = "flip"

        # Restart on a cycle (a point seen before) and periodically.
        h_candidate = integer_key(candidate, int_idx)
        cycled = h_candidate in seen
        periodic = since_restart >= restart_period
        if cycled or periodic:
            perturb(candidate, x, milp, rng, general, is_general, ladder, iteration)
            stats.perturbations += 1
            since_restart = 0
            for _ in range(10):              # until the point is new
                h = integer_key(candidate, int_idx)
                if not (h in seen or h == h_current):
                    break
                perturb(candidate, x, milp, rng, general, is_general, ladder, iteration)
            action = "cycle" if cycled else "restart"
        seen.add(h_current)
        h_current = integer_key(candidate, int_idx) if (cycled or periodic) else h_candidate
        xt = candidate
        if iteration <= 10 or iteration % 10 == 0:
            _print_iteration(emit, iteration, distance, violation, inner, gap_used, action,
                             elapsed())
    found = max_violation(milp, xt) <= tol and is_integral(xt, int_idx, tol)
    return _finish(milp, found, xt, stats, elapsed(), emit, log)


def _print_iteration(emit, iteration, distance, violation, inner, gap, action, t):
    emit(f"{iteration:6d} {distance:12.6g} {violation:10.3g} {inner:10d} {gap:9.2e} "
         f"{action:>8s} {t:8.2f}s\n")


def _finish(milp: MILP, found: bool, xt, stats: _Counters, t: float, emit, log) -> PumpResult:
    """src/pump.jl:412-430."""
    violation = max_violation(milp, xt)
    objective = float(np.dot(milp.c, xt)) + milp.c0
    if found:
        emit(f"FEASIBLE    objective {objective:.10g} after {stats.iterations} iterations, "
             f"{t:.2f} s\n")
    else:
        emit(f"NOT FOUND   after {stats.iterations} iterations, {t:.2f} s "
             f"(max violation {violation:.3g})\n")
    emit(f"            root LP {stats.root_lp_time:.2f} s, projections "
         f"{stats.projection_time:.2f} s, rounding {stats.rounding_time:.2f} s\n")
    return PumpResult(found, np.asarray(xt), objective, violation, stats.iterations,
                      stats.perturbations, stats.pdhg_iterations, stats.root_lp_status,
                      stats.root_lp_time, stats.projection_time, stats.rounding_time, t, log)
