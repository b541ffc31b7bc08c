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
import os
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .milp import (MILP, DeviceViolationCheck, is_integral, max_violation, num_constraints,
                   num_continuous, num_variables)
from .rounding import BatchPolicy, Rounder, jround

# Smallest relative KKT tolerance handed to the FP32 projector form.
FP32_TOLERANCE_FLOOR = 1e-5


def _fp64_successor(P, pp, mixed: bool, device: int):
    """An FP64 projector context that carries on from FP32 projections (the
    polish, or the FP64 form below the FP32 tolerance floor): on an instance
    that fits in LDS its blocks and its power iteration run in one workgroup
    (a user's "fp64" pump keeps the reference's launch structure).  A
    substituted projector without kernel paths is used as it comes."""
    st = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                     precision="fp64", device=device)
    if hasattr(st, "set_kernel_path"):
        st.set_kernel_path("one-workgroup")
    if mixed:
        P.original_primal_residual(st, pp.D_r)
    _column_scale(P, st, pp)
    return st


def _column_scale(P, st, pp) -> None:
    """Hand the projector D_c where it divides targets on the device."""
    if hasattr(P, "set_column_scale"):
        P.set_column_scale(st, pp.D_c)


def _put_target(P, st, pp, xt) -> None:
    """set_target!(projector, x̃ ./ D_c)  (src/pump.jl:331): x̃ in original
    units, divided on the device where the projector does it (bit for bit the
    host's division), on the host otherwise."""
    if hasattr(P, "set_target_original"):
        P.set_target_original(st, xt)
    else:
        P.set_target(st, xt / pp.D_c)


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
    trajectory: int = 0     # feasibility_pump_batch: the trajectory that found x
    seed: int = 0


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
    idx = np.asarray(idx, dtype=np.int64)
    return bool(np.any(np.asarray(a)[idx] != np.asarray(b)[idx]))


def _distance(x, xt, idx) -> float:
    """Σ_{j∈I} |x_j - x̃_j| accumulated in index order, as the reference's
    loop (np.cumsum adds sequentially)."""
    if len(idx) == 0:
        return 0.0
    d = np.abs(np.asarray(x)[idx] - np.asarray(xt)[idx])
    return float(np.cumsum(d)[-1])


def round_to_nearest(x, milp: MILP) -> np.ndarray:
    """x with every integer coordinate replaced by its nearest integer,
    clipped to the column's bounds (src/pump.jl:53-62); each coordinate is
    independent, so the numpy form is exact."""
    xt = np.array(x, dtype=np.float64)
    I = milp.int_idx
    v = np.rint(xt[I])
    lo, hi = milp.l[I], milp.u[I]
    v = np.where(np.isfinite(lo) & (v < lo), np.ceil(lo), v)
    v = np.where(np.isfinite(hi) & (v > hi), np.floor(hi), v)
    xt[I] = v
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
    """Sum, over rows that touch no continuous column, of how far x breaks
    the row (src/pump.jl:98-118)."""
    A = milp.A
    m = num_constraints(milp)
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    activity = np.bincount(A.indices, weights=A.data * np.asarray(x)[cols], minlength=m)
    is_int = milp.is_int
    skip = np.zeros(m, dtype=bool)
    cont_entries = ~is_int[cols]
    skip[A.indices[cont_entries]] = True
    le = milp.row_sense == "le"
    d = activity - milp.b
    viol = np.where(le, np.maximum(d, 0.0), np.abs(d))[~skip]
    # the reference's left-to-right sum over the rows (cumsum adds in order)
    return float(np.cumsum(viol)[-1]) if len(viol) else 0.0


def flip_farthest(xt: np.ndarray, x, milp: MILP, TT: int) -> None:
    """When rounding returns the previous target, move the TT integer
    coordinates of x̃ that lie farthest from x one unit towards x, inside the
    bounds (src/pump.jl:121-135)."""
    l, u = milp.l, milp.u
    I = milp.int_idx
    dist = np.abs(np.asarray(x)[I] - xt[I])
    # descending, ties in index order: sort!(rev = true) is stable
    order = np.argsort(-dist, kind="stable")
    for k in range(min(TT, len(order))):
        d, j = float(dist[order[k]]), int(I[order[k]])
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
    # the binaries: one draw each, in index order (Generator.random(k) gives
    # the k draws of k calls); every update is element-wise
    J = milp.int_idx[~is_general[milp.int_idx]]
    if len(J):
        rho = -0.3 + rng.random(len(J))
        xJ, cur = np.asarray(x)[J], xt[J]
        flip = np.abs(xJ - cur) + np.maximum(rho, 0.0) > 0.5
        direction = np.where(xJ >= cur, 1.0, -1.0)
        v = cur + direction
        out = (v < l[J]) | (v > u[J])
        v = np.where(out, cur - direction, v)
        ok = flip & ~((v < l[J]) | (v > u[J]))
        xt[J[ok]] = v[ok]


def feasibility_pump(milp: MILP, *, time_limit: float = 1800.0, seed: int = 1,
                     verbose: bool = False, flip_parameter: int = 20,
                     restart_period: int = 100, swap_iteration: int = 10,
                     gap_start: float = 1e-2, gap_decay: float = 0.98, gap_floor: float = 1e-4,
                     residual_ratio: float = 0.1, root_lp_tolerance: float = 1e-2,
                     projection_iteration_limit: int = 500_000, batch_policy: str = "cost",
                     push_threshold: float = 0.01, violation_threshold: float = 5000.0,
                     propagation_budget: int = 100_000, feasibility_tolerance: float = 1e-6,
                     precision: str = "fp32", device: int = 0, scaling: str = "device",
                     polish: bool = True, projector=None, root_lp=None) -> PumpResult:
    """Run the feasibility pump on ``milp`` until a feasible point turns up or
    the time limit passes (src/pump.jl:231-410).  The keyword arguments and defaults are the
    reference's (README.md:89-108); ``precision`` picks the projector's form
    ("fp32": FP32 data and iterates with FP64 reductions and polish; "fp64");
    ``scaling`` where the projector's matrix is equilibrated ("device" or
    "host"); ``polish`` (FP32 form, on by default): each FP32 projection is
    finished by an FP64 solve started at the FP32 point (x, y and z) with the
    same target and tolerances, and the rounder receives the FP64 point.

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

    cfg = _Config(time_limit, flip_parameter, restart_period, swap_iteration, gap_start,
                  gap_decay, gap_floor, residual_ratio, projection_iteration_limit,
                  batch_policy, push_threshold, violation_threshold, propagation_budget,
                  feasibility_tolerance, _pdfp_mod.SolverParameters, t0)
    # the device pump checks its points on the device (the original A stays
    # in HBM); with substituted solvers the check runs on the host
    cfg.violation = (DeviceViolationCheck(milp, device) if projector is None
                     else (lambda v: max_violation(milp, v)))
    mixed = num_continuous(milp) > 0
    stats = _Counters()
    log: list = []
    emit = _emitter(log, verbose)
    x, dual_slack, done = _root(milp, cfg, solve_lp, root_lp_tolerance, stats, emit)
    if done is not None:
        return _finish(milp, True, done, stats, elapsed(), emit, log)

    # The projector, built at the first projection: scaled data on the
    # device, warm-started at the root LP solution (src/pump.jl:293-309).
    proj: dict = {}
    m = num_constraints(milp)
    primal_weight = math.sqrt(1.0 + (len(milp.int_idx) / m if m > 0 else 0.0))

    def build():
        pp = _pdfp_mod.projector_problem(milp, scaling=scaling, device=device)
        st = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                         precision=precision, device=device)
        if mixed:
            P.original_primal_residual(st, pp.D_r)
        _column_scale(P, st, pp)
        st.set_x(x / pp.D_c)
        proj.update(pp=pp, state=st)

    def solve_projection(kind, xt, params):
        if not proj:
            build()
        pp = proj["pp"]
        # x̃ in original units; the trajectory leaves it unchanged until its
        # next projection, which brings a new one
        target = None if kind == "tighten" else xt
        # FP32 iterates do not resolve relative KKT errors much below 1e-5:
        # from the first projection asked for less, the FP64 form takes over,
        # started at the last FP32 projection.
        st = proj["state"]
        if (getattr(st, "precision", None) == "fp32" and
                min(params.eps_optimal_relative,
                    params.eps_feasible_relative) < FP32_TOLERANCE_FLOOR):
            st64 = _fp64_successor(P, pp, mixed, device)
            st64.set_x(st.pdhg_x)
            _put_target(P, st64, pp, proj["target"] if target is None else target)
            emit(f"projector   FP64 form from gap tolerance "
                 f"{params.eps_optimal_relative:.2e} on\n")
            proj["state"] = st = st64
        elif target is not None:
            _put_target(P, st, pp, target)
        else:
            P.warm_start(st)
        if target is not None:
            proj["target"] = target
        P.pdfp_solve(st, params, primal_weight)
        if polish and getattr(st, "precision", None) == "fp32":
            # the FP64 finish: started at the FP32 point (current and initial
            # x, y, z), the same target and tolerances
            st64 = proj.get("polish")
            if st64 is None:
                st64 = _fp64_successor(P, pp, mixed, device)
                proj["polish"] = st64
            st64.set_x(st.pdhg_x)
            st64.set_dual(st.vector("pdhg_y"), st.vector("pdhg_z"))
            _put_target(P, st64, pp, proj["target"])
            P.pdfp_solve(st64, params, primal_weight)
            return pp.D_c * st64.pdhg_x, st.total_count + st64.total_count
        return pp.D_c * st.pdhg_x, st.total_count

    traj = _trajectory(milp, cfg, seed, x, dual_slack, stats, emit)
    found, xt = _drive(traj, solve_projection)
    return _finish(milp, found, xt, stats, elapsed(), emit, log)


@dataclass
class _Config:
    time_limit: float
    flip_parameter: int
    restart_period: int
    swap_iteration: int
    gap_start: float
    gap_decay: float
    gap_floor: float
    residual_ratio: float
    projection_iteration_limit: int
    batch_policy: str
    push_threshold: float
    violation_threshold: float
    propagation_budget: int
    feasibility_tolerance: float
    params_cls: type
    t0: float
    violation: object = None   # x -> max_violation(milp, x)
    stop: object = None        # a threading.Event: set, the trajectories end at their next step

    def stopped(self) -> bool:
        return self.stop is not None and self.stop.is_set()

    def elapsed(self) -> float:
        return time.perf_counter() - self.t0

    def remaining(self) -> float:
        return max(self.time_limit - self.elapsed(), 1.0)


def _emitter(log: list, verbose: bool):
    def emit(s: str):
        log.append(s)
        if verbose:
            sys.stdout.write(s)
            sys.stdout.flush()
    return emit


def _root(milp, cfg: _Config, solve_lp, root_lp_tolerance, stats, emit):
    """The header and the root LP relaxation (src/pump.jl:245-279); returns
    (x, dual_slack, the rounded point when it is already feasible)."""
    n = num_variables(milp)
    m = num_constraints(milp)
    mixed = num_continuous(milp) > 0
    emit(f"PDFP feasibility pump: {milp.name}\n")
    emit(f"model       {m} rows, {n} columns ({len(milp.int_idx)} integer, "
         f"{num_continuous(milp)} continuous), {milp.A.nnz} nonzeros\n")
    x, dual_slack, stats.root_lp_status, root_iterations = solve_lp(
        milp, tolerance=root_lp_tolerance, time_limit=cfg.remaining(),
        cupdlpx_termination=mixed)
    stats.root_lp_time = cfg.elapsed()
    emit(f"root LP     {stats.root_lp_status} after {root_iterations} iterations, "
         f"{stats.root_lp_time:.2f} s\n")
    if is_integral(x, milp.int_idx, cfg.feasibility_tolerance):
        xt = round_to_nearest(x, milp)
        if cfg.violation(xt) <= cfg.feasibility_tolerance:
            return x, dual_slack, xt
    return x, dual_slack, None


def _drive(traj, solve):
    """Run one trajectory generator to its end with a projection callable."""
    try:
        req = traj.send(None)
        while True:
            req = traj.send(solve(*req))
    except StopIteration as e:
        return e.value


def _trajectory(milp, cfg: _Config, seed: int, x, dual_slack, stats, emit):
    """One pump trajectory from the first rounding on (src/pump.jl:281-410),
    as a generator: it yields each projection it needs, ("target", x̃,
    params) for x̃ in original units or ("tighten", None, params) for a
    warm-started re-solve of the same target, receives (x in original
    units, PDHG iterations), and returns (found, x̃)."""
    n = num_variables(milp)
    int_idx = milp.int_idx
    int_list = int_idx.tolist()
    tol = cfg.feasibility_tolerance
    mixed = num_continuous(milp) > 0
    rng = np.random.Generator(np.random.PCG64(seed))
    rounder = Rounder(milp, max_pops=(2**63 - 1) if mixed else cfg.propagation_budget,
                      swap_iteration=cfg.swap_iteration,
                      policy=BatchPolicy() if cfg.batch_policy == "cost" else None)
    params_cls = cfg.params_cls

    # Round the root LP point once; a large row violation of that rounding switches on the guided rule.
    rounder.schedule = batch_schedule(milp, x, dual_slack, cfg.push_threshold)
    t = time.perf_counter()
    xt = rounder.round_guarded(x, rng, 0)
    stats.rounding_time += time.perf_counter() - t
    total_violation = total_integer_row_violation(milp, xt)
    if total_violation >= cfg.violation_threshold:
        direction = [False] * n
        for j in int_list:
            direction[j] = True
        rounder.direction = direction
    emit(f"rounder     {rounder.schedule} schedule, violation-guided rounding "
         f"{'off' if rounder.direction is None else 'on'} (first rounding violates by "
         f"{total_violation:.4g})\n")

    gap = cfg.gap_start
    tightening_budget = cfg.gap_start
    general = [j for j in int_list if not (milp.l[j] == 0.0 and milp.u[j] == 1.0)]
    is_general = np.zeros(n, dtype=bool)
    is_general[general] = True
    ladder = RestartLadder(0.0, 0, cfg.flip_parameter, max(len(general) // 10, 10))
    seen: set = set()
    h_current = integer_key(xt, int_idx)
    since_restart = 0

    emit(f"{'iter':>6s} {'distance':>12s} {'violation':>10s} {'pdhg':>10s} {'gap tol':>9s} "
         f"{'action':>8s} {'time':>9s}\n")
    while cfg.elapsed() < cfg.time_limit and not cfg.stopped():
        violation = cfg.violation(xt)
        if violation <= tol and is_integral(xt, int_idx, tol):
            return True, xt.copy()
        stats.iterations += 1
        iteration = stats.iterations
        since_restart += 1

        # ℓ1 projection of the current integer target onto the LP polytope.
        t = time.perf_counter()
        gap_used = gap
        params = params_cls(iteration_limit=cfg.projection_iteration_limit,
                            eps_optimal_relative=gap,
                            eps_feasible_relative=gap * cfg.residual_ratio,
                            time_limit=cfg.remaining())
        x, inner = yield ("target", xt, params)
        if cfg.stopped():
            return False, xt
        # Mixed models: when the projection is integral but rounding it still
        # breaks a row (so the continuous part is off), solve again at a
        # tenfold tighter gap, at most five times a run (src/pump.jl:340-356).
        while (mixed and tightening_budget > 1e-6 and cfg.elapsed() < cfg.time_limit
               and is_integral(x, int_idx, tol)
               and cfg.violation(round_to_nearest(x, milp)) > tol):
            tightening_budget *= 0.1
            gap = max(gap * 0.1, 1e-8)
            params = params_cls(iteration_limit=cfg.projection_iteration_limit,
                                eps_optimal_relative=gap,
                                eps_feasible_relative=gap * cfg.residual_ratio,
                                time_limit=cfg.remaining())
            x, count = yield ("tighten", None, params)
            inner += count
            if cfg.stopped():
                return False, xt
        stats.pdhg_iterations += inner
        if gap > cfg.gap_floor:
            gap = max(gap * cfg.gap_decay, cfg.gap_floor)
        stats.projection_time += time.perf_counter() - t
        distance = _distance(x, xt, int_idx)

        if is_integral(x, int_idx, tol):
            candidate = round_to_nearest(x, milp)
            if cfg.violation(candidate) <= tol:
                _print_iteration(emit, iteration, distance, violation, inner, gap_used,
                                 "integral", cfg.elapsed())
                return True, candidate

        # New target from the projection; if rounding reproduces the old target, flip coordinates instead.
        t = time.perf_counter()
        x_round = rounder.round_guarded(x, rng, iteration)
        stats.rounding_time += time.perf_counter() - t
        candidate = xt.copy()
        action = "round"
        if _differs(x_round, xt, int_idx):
            candidate[:] = x_round
        else:
            flips = int(rng.integers(cfg.flip_parameter // 2, 3 * cfg.flip_parameter // 2 + 1))
            flip_farthest(candidate, x, milp, flips)
            action = "flip"

        # Perturb when the target repeats an earlier one, and every restart_period iterations.
        h_candidate = integer_key(candidate, int_idx)
        cycled = h_candidate in seen
        periodic = since_restart >= cfg.restart_period
        if cycled or periodic:
            perturb(candidate, x, milp, rng, general, is_general, ladder, iteration)
            stats.perturbations += 1
            since_restart = 0
            for _ in range(10):              # at most 10 retries for an unseen target
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
                             cfg.elapsed())
    found = cfg.violation(xt) <= tol and is_integral(xt, int_idx, tol)
    return found, xt


def feasibility_pump_batch(milp: MILP, *, trajectories: int = 8, seed: int = 1,
                           time_limit: float = 1800.0, verbose: bool = False,
                           flip_parameter: int = 20, restart_period: int = 100,
                           swap_iteration: int = 10, gap_start: float = 1e-2,
                           gap_decay: float = 0.98, gap_floor: float = 1e-4,
                           residual_ratio: float = 0.1, root_lp_tolerance: float = 1e-2,
                           projection_iteration_limit: int = 500_000,
                           batch_policy: str = "cost", push_threshold: float = 0.01,
                           violation_threshold: float = 5000.0,
                           propagation_budget: int = 100_000,
                           feasibility_tolerance: float = 1e-6, device: int = 0,
                           scaling: str = "device", polish: bool = True) -> PumpResult:
    """K concurrent pump trajectories on one GPU; the first feasible point
    wins.  Replaces solve_concurrent (scripts/solve.jl:121-148), which forks K
    processes with seeds 100k + seed, each holding its own copy of the
    instance on the GPU: here the trajectories (seeds 100k + seed, k = 1..K)
    share one root LP and one batched FP32 projector (PDFPBatch, the matrix
    in HBM once).  The projector runs continuously: a trajectory whose
    projection ends is rounded on a host thread while the others keep
    iterating, and re-enters the batch with its next target at the next
    evaluation block.  A trajectory asked for a tolerance tighter than the
    FP32 form resolves moves to an FP64 projector of its own, as in
    feasibility_pump.  The result is the winning trajectory's, with
    ``trajectory`` and ``seed`` set; its ``time`` is the time to the first
    feasible point of the whole run."""
    if not 1 <= trajectories <= 16:
        raise ValueError("trajectories must be in 1..16")
    if batch_policy not in ("cost", "always"):
        raise ValueError("batch_policy must be 'cost' or 'always'")
    from . import pdfp as _pdfp_mod
    from . import pdlp as _pdlp_mod
    from .batch import PDFPBatch

    K = trajectories
    t0 = time.perf_counter()
    cfg = _Config(time_limit, flip_parameter, restart_period, swap_iteration, gap_start,
                  gap_decay, gap_floor, residual_ratio, projection_iteration_limit,
                  batch_policy, push_threshold, violation_threshold, propagation_budget,
                  feasibility_tolerance, _pdfp_mod.SolverParameters, t0)
    checker = DeviceViolationCheck(milp, device)
    lock = threading.Lock()

    def violation(v):
        with lock:              # one checker, trajectories on several threads
            return checker(v)

    cfg.violation = violation
    cfg.stop = threading.Event()
    mixed = num_continuous(milp) > 0
    root_stats = _Counters()
    root_log: list = []

    def solve_lp(milp_, **kw):
        return _pdlp_mod.solve_lp_relaxation(milp_, scaling=scaling, device=device, **kw)

    x, dual_slack, done = _root(milp, cfg, solve_lp, root_lp_tolerance, root_stats,
                                _emitter(root_log, verbose))
    seeds = [100 * (k + 1) + seed for k in range(K)]
    if done is not None:
        r = _finish(milp, True, done, root_stats, cfg.elapsed(), _emitter(root_log, verbose),
                    root_log)
        r.seed = seeds[0]
        return r

    stats = []
    logs = []
    gens = []
    for k in range(K):
        st = _Counters(root_lp_status=root_stats.root_lp_status,
                       root_lp_time=root_stats.root_lp_time)
        lg = list(root_log)
        lg.append(f"trajectory  {k + 1} of {K}, seed {seeds[k]}\n")
        stats.append(st)
        logs.append(lg)
        gens.append(_trajectory(milp, cfg, seeds[k], x, dual_slack, st,
                                _emitter(lg, verbose and k == 0)))

    pp = _pdfp_mod.projector_problem(milp, scaling=scaling, device=device)
    D_c = pp.D_c
    batch = PDFPBatch(K, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                      device=device)
    if mixed:
        batch.original_primal_residual(pp.D_r)
    batch.set_column_scale(D_c)
    for k in range(K):
        batch.set_x(k, x / D_c)
    m = num_constraints(milp)
    primal_weight = math.sqrt(1.0 + (len(milp.int_idx) / m if m > 0 else 0.0))
    fp64: list = [None] * K        # a trajectory's own FP64 projector, once it needs one
    target: list = [None] * K      # its last target (original units)
    fin64: list = [None] * K       # its FP64 context for the polish of its FP32 projections

    def polished(k, x_s, y, z, count, params):
        """Trajectory k's FP32 projection finished by an FP64 solve from its
        point (x, y, z), on the host thread that then rounds it."""
        st = fin64[k]
        if st is None:
            st = _fp64_successor(_pdfp_mod, pp, mixed, device)
            fin64[k] = st
        st.set_x(x_s)
        st.set_dual(y, z)
        _put_target(_pdfp_mod, st, pp, target[k])
        _pdfp_mod.pdfp_solve(st, params, primal_weight)
        return advance(k, (D_c * st.pdhg_x, count + st.total_count))

    def finish(k, found, xt, t=None):
        r = _finish(milp, found, xt, stats[k], cfg.elapsed() if t is None else t,
                    _emitter(logs[k], False), logs[k])
        r.trajectory = k
        r.seed = seeds[k]
        return r

    pending = {}
    for k in range(K):
        try:
            pending[k] = gens[k].send(None)
        except StopIteration as e:
            found, xt = e.value
            if found:
                return finish(k, True, xt)
    last = None
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait

    # Continuous batching: a trajectory whose projection ends is handed to a
    # host thread (rounding, flips, restarts) while the device runs on with
    # the others, and re-enters the batch with its next projection at the
    # next block boundary.
    pool = ThreadPoolExecutor(max_workers=max(1, min(K, os.cpu_count() or 1)))

    def advance(k, result):
        try:
            return gens[k].send(result), None
        except StopIteration as e:
            return None, e.value

    running: set = set()
    futures: dict = {}
    batch_params: dict = {}
    winner = None
    while pending or running or futures:
        for k in sorted(pending):
            kind, xt, params = pending[k]
            if kind == "target":
                target[k] = xt  # unchanged by its trajectory until the next projection
            tight = min(params.eps_optimal_relative,
                        params.eps_feasible_relative) < FP32_TOLERANCE_FLOOR
            if fp64[k] is None and tight:
                st64 = _fp64_successor(_pdfp_mod, pp, mixed, device)
                st64.set_x(batch.pdhg_x(k))
                _put_target(_pdfp_mod, st64, pp, target[k])
                fp64[k] = st64
                kind = "started"
            if fp64[k] is not None:
                st64 = fp64[k]
                if kind == "target":
                    _put_target(_pdfp_mod, st64, pp, target[k])
                elif kind == "tighten":
                    _pdfp_mod.warm_start(st64)
                _pdfp_mod.pdfp_solve(st64, params, primal_weight)
                futures[k] = pool.submit(advance, k, (D_c * st64.pdhg_x, st64.total_count))
            else:
                if kind == "target":
                    batch.set_target_original(k, target[k])
                else:
                    batch.warm_start(k)
                batch.begin(k, params, primal_weight)
                running.add(k)
                batch_params[k] = params
        pending = {}
        if running:
            # with projections waiting on the host, return after each block
            # so that those that end their rounding re-enter at the next one
            terms, _ = batch.run(until_any=True, one_block=bool(futures))
            for k in sorted(running):
                if terms[k] != "UNSPECIFIED":
                    running.discard(k)
                    if polish:
                        futures[k] = pool.submit(polished, k, batch.pdhg_x(k),
                                                 batch.vector(k, "pdhg_y"),
                                                 batch.vector(k, "pdhg_z"),
                                                 batch.total_count(k), batch_params[k])
                    else:
                        futures[k] = pool.submit(advance, k, (D_c * batch.pdhg_x(k),
                                                              batch.total_count(k)))
        if futures:
            done = [k for k, f in futures.items() if f.done()]
            if not done and not running:
                wait(list(futures.values()), return_when=FIRST_COMPLETED)
                done = [k for k, f in futures.items() if f.done()]
            for k in sorted(done):
                req, end = futures.pop(k).result()
                if end is None:
                    pending[k] = req
                elif end[0]:
                    winner = winner if winner is not None else (k, end[1])
                else:
                    last = (k, end[1])
        if winner is not None:
            # the time to the first feasible point is taken now; the other
            # trajectories' host work in flight ends at its next step, and is
            # waited for, so that none of it outlives this call
            t_win = cfg.elapsed()
            cfg.stop.set()
            pool.shutdown(wait=True, cancel_futures=True)
            return finish(winner[0], True, winner[1], t_win)
    pool.shutdown()
    k, xt = last if last is not None else (0, x)
    return finish(k, False, xt)


def feasibility_pump_pack(milps, *, seed: int = 1, time_limit: float = 1800.0,
                          verbose: bool = False, flip_parameter: int = 20,
                          restart_period: int = 100, swap_iteration: int = 10,
                          gap_start: float = 1e-2, gap_decay: float = 0.98,
                          gap_floor: float = 1e-4, residual_ratio: float = 0.1,
                          root_lp_tolerance: float = 1e-2,
                          projection_iteration_limit: int = 500_000,
                          batch_policy: str = "cost", push_threshold: float = 0.01,
                          violation_threshold: float = 5000.0,
                          propagation_budget: int = 100_000,
                          feasibility_tolerance: float = 1e-6, device: int = 0,
                          scaling: str = "device") -> list:
    """The pumps of K independent instances on one GPU, one trajectory each
    (``feasibility_pump``'s, seed ``seed``): the root LPs are solved one by
    one, then every instance's projections run in one block-diagonal pack
    (PDFPPack) that iterates continuously: an instance whose projection
    ends is rounded on a host thread while the others keep iterating, and
    re-enters the pack with its next target at the next evaluation block.
    An instance asked for a tolerance tighter than the FP32 form resolves
    moves to an FP64 projector of its own.  Returns one PumpResult per
    instance; each ``time`` counts from the start of the whole run."""
    if batch_policy not in ("cost", "always"):
        raise ValueError("batch_policy must be 'cost' or 'always'")
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait

    from . import pdfp as _pdfp_mod
    from . import pdlp as _pdlp_mod
    from .pack import PDFPPack

    K = len(milps)
    t0 = time.perf_counter()

    def solve_lp(milp_, **kw):
        return _pdlp_mod.solve_lp_relaxation(milp_, scaling=scaling, device=device, **kw)

    cfgs, stats, logs, gens = [], [], [], [None] * K
    results: list = [None] * K
    starts = []
    for k, milp in enumerate(milps):
        cfg = _Config(time_limit, flip_parameter, restart_period, swap_iteration, gap_start,
                      gap_decay, gap_floor, residual_ratio, projection_iteration_limit,
                      batch_policy, push_threshold, violation_threshold,
                      propagation_budget, feasibility_tolerance,
                      _pdfp_mod.SolverParameters, t0)
        cfg.violation = DeviceViolationCheck(milp, device)
        st = _Counters()
        lg: list = []
        emit = _emitter(lg, verbose and k == 0)
        x, dual_slack, done = _root(milp, cfg, solve_lp, root_lp_tolerance, st, emit)
        cfgs.append(cfg)
        stats.append(st)
        logs.append(lg)
        starts.append(x)
        if done is not None:
            results[k] = _finish(milp, True, done, st, cfg.elapsed(), emit, lg)
        else:
            gens[k] = _trajectory(milp, cfg, seed, x, dual_slack, st, emit)

    live = [k for k in range(K) if gens[k] is not None]
    if not live:
        return results
    pps = [_pdfp_mod.projector_problem(milps[k], scaling=scaling, device=device)
           for k in range(K)]
    pack = PDFPPack(pps, device=device)
    mixed = [num_continuous(m) > 0 for m in milps]
    for k in range(K):
        if mixed[k]:
            pack.original_primal_residual(k, pps[k].D_r)
        pack.set_x(k, starts[k] / pps[k].D_c)
    pw = [math.sqrt(1.0 + (len(m.int_idx) / num_constraints(m)
                           if num_constraints(m) > 0 else 0.0)) for m in milps]
    fp64: list = [None] * K
    target: list = [None] * K
    last: list = [None] * K

    def finish(k, found, xt):
        results[k] = _finish(milps[k], found, xt, stats[k], cfgs[k].elapsed(),
                             _emitter(logs[k], False), logs[k])

    def advance(k, result):
        try:
            return gens[k].send(result), None
        except StopIteration as e:
            return None, e.value

    pending = {}
    for k in live:
        try:
            pending[k] = gens[k].send(None)
        except StopIteration as e:
            finish(k, e.value[0], e.value[1])
    pool = ThreadPoolExecutor(max_workers=max(1, min(len(live), os.cpu_count() or 1)))
    running: set = set()
    futures: dict = {}
    while pending or running or futures:
        for k in sorted(pending):
            kind, xt, params = pending[k]
            pp = pps[k]
            if kind == "target":
                target[k] = xt / pp.D_c
            tight = min(params.eps_optimal_relative,
                        params.eps_feasible_relative) < FP32_TOLERANCE_FLOOR
            if fp64[k] is None and tight:
                st64 = _fp64_successor(_pdfp_mod, pp, mixed[k], device)
                st64.set_x(pack.pdhg_x(k))
                _pdfp_mod.set_target(st64, target[k])
                fp64[k] = st64
                kind = "started"
            if fp64[k] is not None:
                st64 = fp64[k]
                if kind == "target":
                    _pdfp_mod.set_target(st64, target[k])
                elif kind == "tighten":
                    _pdfp_mod.warm_start(st64)
                _pdfp_mod.pdfp_solve(st64, params, pw[k])
                futures[k] = pool.submit(advance, k, (pp.D_c * st64.pdhg_x,
                                                      st64.total_count))
            else:
                if kind == "target":
                    pack.set_target(k, target[k])
                else:
                    pack.warm_start(k)
                pack.begin(k, params, pw[k])
                running.add(k)
        pending = {}
        if running:
            # with projections waiting on the host, return after each block
            # so that those that end their rounding re-enter at the next one
            terms, _ = pack.run(until_any=True, one_block=bool(futures))
            for k in sorted(running):
                if terms[k] != "UNSPECIFIED":
                    running.discard(k)
                    futures[k] = pool.submit(advance, k, (pps[k].D_c * pack.pdhg_x(k),
                                                          pack.total_count(k)))
        if futures:
            done = [k for k, f in futures.items() if f.done()]
            if not done and not running:
                wait(list(futures.values()), return_when=FIRST_COMPLETED)
                done = [k for k, f in futures.items() if f.done()]
            for k in sorted(done):
                req, end = futures.pop(k).result()
                if end is None:
                    pending[k] = req
                else:
                    finish(k, end[0], end[1])
    pool.shutdown()
    for k in range(K):
        if results[k] is None:        # a generator that never ran to an end
            finish(k, False, last[k] if last[k] is not None else starts[k])
    return results


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
