"""Benchmark of the PDHG projector on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload powerlaw-5m|mixed-1m|setcover-pack-64|general-int-4m]

Workload (default, "powerlaw-5m"): the mixed MILP with 500,000 rows × 1,000,000
columns (70 % binary, 30 % continuous in [0, 10]) and exactly 5,000,000
nonzeros in power-law row lengths (Zipf 1.5, longest rows 5,000), values
N(0,1), b = A x0 + |N(0,1)| for a random feasible x0 (seed 1): the FP32-bulk /
FP64-residual configuration on which the targets (batched FP32 at 3x the plain
FP64 form, the SpMV against the HBM roofline) are stated.  ``--workload
mixed-1m``: the same generator at 50,000 × 100,000 with 1M nonzeros in rows of
10-30; ``general-int-4m``: a general-integer, equality-heavy MILP of 200,000 ×
400,000 with 4M nonzeros; ``setcover-pack-64``: 64 set-covering instances of
uneven sizes in one block-diagonal pack (15.5M nonzeros in all).  A step is one warm-started ℓ1
projection toward a new random integer target (0/1 on binaries) at relative
KKT tolerance 1e-4, the way the pump calls the projector
(src/pump.jl:329-359).  The scaled matrix, its transpose and every
iterate stay in HBM; the timed region covers the pump's set_target!(x̃ ./ D_c)
(x̃ in original units, the division on the device) + pdfp_solve.

value = Σ PDHG iterations / projection wall time of the batched product
form: 8 trajectory slots sharing the matrix in HBM, FP32 data and iterates,
fused launches, FP64 reductions, serving the step's projections as a queue
(a slot that finishes takes the next target while the others run on;
``--batch 1``: a single trajectory), the reference's own accounting
(src/pump.jl:357-359).  The single-trajectory
FP32 form and the plain FP64 form (the reference's launch structure) run the
same kind of projections and are reported beside it.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch over its mean duration, timed with the kernel's own start/stop events
on the library's stream; HBM traffic per launch from the committed rocprofv3
counter summary of the same workload, when there is one), the CPU baseline
(the FP64 C oracle on one host core, on a bounded number of iterations of the
same workload) and the time to the first feasible point of four MILPs and of the
four example instances' pumps on one pack.

    python bench.py --pump --workload general-int-4m [--pump-time-limit 120]

runs the pump to its first feasible point on a workload's MILP (FP32 with the
FP64 polish, then 8 trajectories on the batch) and prints its times.

    python bench.py --copy-calibration

copies 1 GiB device to device ten times (the plain copy kernel whose
FETCH_SIZE / WRITE_SIZE counters calibrate the traffic figures).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PDHG iter/s + time-to-first-feasible, batched FP32 vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0


def workload(name: str):
    from pdfp_amd import generators as G

    if name == "mixed-1m":
        m = G.mixed(50_000, 100_000, 1)
        desc = ("mixed MILP 50000 rows x 100000 cols (70% binary, 30% continuous [0,10]), "
                "1M nnz, rows of 10-30, N(0,1) values, b = A x0 + |N(0,1)|, seed 1")
    elif name == "powerlaw-5m":
        m = G.mixed(500_000, 1_000_000, 1, power_law=True, nnz=5_000_000, max_len=5000)
        desc = ("mixed MILP 500000 rows x 1000000 cols, 5M nnz, power-law row lengths "
                "(Zipf 1.5, max 5000), N(0,1) values, seed 1")
    elif name == "general-int-4m":
        m = G.general_integer(200_000, 400_000, 1)
        desc = ("general-integer MILP 200000 rows (40% equality) x 400000 integer cols in "
                "[0, U_j], U_j in 1..20, 4M nnz of U{-9..9}, rows of 10-30, seed 1")
    else:
        raise SystemExit(f"unknown workload {name}")
    return m, desc


def kernel_bytes_fp32(m: int, n: int, nnz: int) -> tuple[float, float]:
    """Algorithmic HBM bytes of the two fused FP32 launches of one iteration.
    Transpose pass (columns of A): Aᵀ values+indices 8·nnz, row pointers
    4(n+1), z read once 4m, read x, x⁰, y, y⁰, ℓ, u, x̃, w 32n, write x, y,
    x̄ 12n.  A pass (rows): 8·nnz, 4(m+1), x̄ read once 4n, read z, z⁰, b 12m,
    is_eq m, write z 4m."""
    at = 8 * nnz + 4 * (n + 1) + 4 * m + 32 * n + 12 * n
    a = 8 * nnz + 4 * (m + 1) + 4 * n + 12 * m + m + 4 * m
    return float(at), float(a)


def kernel_bytes_fp64(m: int, n: int, nnz: int) -> tuple[float, float]:
    """Algorithmic HBM bytes of the two SpMVs of the plain FP64 form:
    12·nnz + 4(rows+1) + 8·(gathered vector) + 8·rows (the output)."""
    at = 12 * nnz + 4 * (n + 1) + 8 * m + 8 * n
    a = 12 * nnz + 4 * (m + 1) + 8 * n + 8 * m
    return float(at), float(a)


def spmv_figures(m: int, n: int, nnz: int, spmv_ms):
    """The standalone FP32 SpMVs y = A x and y = Aᵀ x on the resident CSR
    operators (FP64 accumulation, FP32 output): algorithmic bytes
    8·nnz + 4(rows+1) + 4·cols + 4·rows over the mean launch duration, beside
    the random-gather floor of the same shape (nnz 4-byte gathers from a table
    of the gathered vector's size)."""
    from pdfp_amd._lib import gather_ceiling_us

    out = {}
    for name, ms, rows, cols in (("Ax", spmv_ms[0], m, n), ("ATz", spmv_ms[1], n, m)):
        byt = 8.0 * nnz + 4.0 * (rows + 1) + 4.0 * cols + 4.0 * rows
        gbs = byt / (ms * 1e-3) / 1e9
        ceil = gather_ceiling_us(max(cols, 1) * 4, 4, nnz)
        out[name] = {"us": ms * 1e3, "algorithmic_bytes": byt, "achieved_gbs": gbs,
                     "frac": gbs / HBM_PEAK_GBS, "roofline_us": byt / HBM_PEAK_GBS / 1e3,
                     "gather_ceiling_us": ceil, "frac_of_gather_ceiling": ceil / (ms * 1e3)}
    return out


def run_projections(s, pp, milp, targets, tol):
    from pdfp_amd import pdfp as P

    iters = 0
    t_total = 0.0
    statuses = []
    quality = []   # per projection: ℓ1 distance and FP64-evaluated KKT terms
    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    P.set_column_scale(s, pp.D_c)  # once per run: x̃ ./ D_c then runs on the device
    for xt in targets:
        p = P.SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                               iteration_limit=500_000)
        t0 = time.perf_counter()
        P.set_target_original(s, xt)
        st = P.pdfp_solve(s, p, pw)
        t_total += time.perf_counter() - t0
        iters += s.total_count
        statuses.append(st)
        inf = s.info()
        quality.append({"l1": inf.l1_distance, "rel_primal": inf.relative_primal_residual,
                        "rel_dual": inf.relative_dual_residual, "rel_gap": inf.relative_l1_gap})
    run_projections.quality = quality
    return iters, t_total, statuses


def kernel_bytes_batch(m: int, n: int, nnz: int, K: int = 8) -> tuple[float, float]:
    """Algorithmic bytes of the two batched launches with W = 8 trajectory
    slots per column (K ≤ 8) or W = 16 (K ≤ 16), FP32, r = 4W bytes per
    interleaved record.  Transpose pass: 8·nnz + 4(n+1), z of the W slots
    read once r·m, per column x, y, x⁰, y⁰, x̃ (r bytes each) and ℓ, u, w
    (4 B each) (5r + 12)·n, write x, y, x̄ 3r·n.  A pass: 8·nnz + 4(m+1),
    x̄ read once r·n, per row z, z⁰ (r bytes each), b, is_eq (2r + 5)·m,
    write z r·m."""
    r = 4 * (8 if K <= 8 else 16)
    at = 8 * nnz + 4 * (n + 1) + r * m + (5 * r + 12) * n + 3 * r * n
    a = 8 * nnz + 4 * (m + 1) + r * n + (2 * r + 5) * m + r * m
    return float(at), float(a)


def slice_rows(A, min_len: int = 0) -> tuple[int, int]:
    """Rows of the stacked column-slice matrix of the batch's sliced A pass
    (the (row, slice) pairs of A's long rows with a nonzero, slices of n/8
    columns) and the number of long rows: rows with more than ``min_len``
    nonzeros (0: every row)."""
    csr = A.tocsr()
    n = csr.shape[1]
    lens = np.diff(csr.indptr)
    long_rows = lens > min_len if min_len > 0 else np.ones(csr.shape[0], dtype=bool)
    rows = np.repeat(np.arange(csr.shape[0], dtype=np.int64), lens)
    keep = np.repeat(long_rows, lens)
    sl = (csr.indices.astype(np.int64) * 8) // max(n, 1)
    return int(np.unique(rows[keep] * 8 + sl[keep]).size), int(long_rows.sum())


def kernel_bytes_sliced_a(m: int, n: int, nnz: int, T: int, K: int = 8, *, min_len: int = 0,
                          nnz_long: int | None = None, n_long: int | None = None) -> float:
    """Algorithmic bytes of the batch's A pass over 8 column slices (the
    layout's own figure).  Every row over the slices (min_len 0): the
    slices' pass 8·nnz + 4(T + 9) (stacked row pointers) + 4T (each stacked
    row's partial position) + x̄ read once r·n + its partials written r·T;
    the combine 8·m·r (the 8 partials of every row) + z, z⁰, b, is_eq
    (2r + 5)·m + write z r·m.  Only the rows longer than min_len over the
    slices: the slices' pass over the long rows' nnz_long nonzeros as above
    (x̄ still counted once), then the one pass over the short rows
    8·(nnz - nnz_long) + 4(m+1) + 4m (each row's long-row index) + the long
    rows' partials 8·n_long·r + the z update's (2r + 5)·m + r·m."""
    r = 4 * (8 if K <= 8 else 16)
    if min_len == 0:
        return float(8 * nnz + 4 * (T + 9) + 4 * T + r * n + r * T + 8 * m * r
                     + (2 * r + 5) * m + r * m)
    p1 = 8 * nnz_long + 4 * (T + 9) + 4 * T + r * n + r * T
    p2 = 8 * (nnz - nnz_long) + 4 * (m + 1) + 4 * m + 8 * n_long * r + (2 * r + 5) * m + r * m
    return float(p1 + p2)


def run_batch(bt, pp, milp, target_sets, tol):
    """The K trajectory slots serve a queue of warm-started projections, the
    targets of all steps in order (step s, slot k): a slot whose projection
    ends takes the next target at the next evaluation block while the others
    run on (pdfp_batch_begin / pdfp_batch_run), so no slot idles while
    another finishes.  Returns (Σ PDHG iterations, seconds, statuses)."""
    from pdfp_amd.pdfp import SolverParameters

    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=500_000)
    queue = [xt for targets in target_sets for xt in targets]
    iters = 0
    statuses = []
    bt.set_column_scale(pp.D_c)  # once per run: x̃ ./ D_c then runs on the device
    t0 = time.perf_counter()
    running = set()
    nxt = 0
    for k in range(min(bt.K, len(queue))):
        bt.set_target_original(k, queue[nxt])
        bt.begin(k, p, pw)
        running.add(k)
        nxt += 1
    while running:
        terms, _ = bt.run(until_any=True)
        for k in sorted(running):
            if terms[k] == "UNSPECIFIED":
                continue
            running.discard(k)
            iters += bt.last_total_count[k]
            statuses.append(terms[k])
            if nxt < len(queue):
                bt.set_target_original(k, queue[nxt])
                bt.begin(k, p, pw)
                running.add(k)
                nxt += 1
    return iters, time.perf_counter() - t0, statuses


def cpu_baseline(pp, milp, target, iterations: int, threads: int = 1):
    """The FP64 C oracle (the reference's CPU path restated) on one core:
    PDHG iterations of one projection of the same workload, ‖A‖₂ excluded.
    threads > 1: the same with each step's products and elementwise updates
    split over that many host threads (bit-identical; not the reference's
    path)."""
    from oracle import oracle as O
    from pdfp_amd.pdfp import SolverParameters

    o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    O.set_target(o, target / pp.D_c)
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=iterations)
    pw = float(np.sqrt(1.0 + len(milp.int_idx) / max(len(milp.b), 1)))
    o.initialize(p, pw)                       # power iteration, untimed
    O.set_threads(threads)
    try:
        t0 = time.perf_counter()
        O.pdfp_solve(o, p, pw)
        dt = time.perf_counter() - t0
    finally:
        O.set_threads(1)
    it = o.total_count
    if threads == 1:
        return {"value": it / dt, "unit": "PDHG iter/s", "cores": 1, "kind": "port",
                "sample": f"{it} PDHG iterations of one projection of the same workload "
                          f"(FP64 C oracle, single thread, ||A||_2 excluded), {dt:.1f} s"}
    return {"value": it / dt, "unit": "PDHG iter/s", "cores": threads, "kind": "port",
            "not_the_reference": True,
            "sample": f"{it} PDHG iterations of one projection of the same workload (FP64 C "
                      f"oracle, each step's products and updates split over {threads} "
                      f"threads, bit-identical to one thread; the reference's CPU path is "
                      f"single-threaded), {dt:.1f} s"}


def _oracle_pump_backends():
    """The oracle as the pump's projector and root LP (CPU baseline leg)."""
    from types import SimpleNamespace

    from oracle import oracle as O

    proj = SimpleNamespace(PDFPState=O.OraclePDFPState, set_target=O.set_target,
                           warm_start=O.warm_start, pdfp_solve=O.pdfp_solve,
                           original_primal_residual=O.original_primal_residual)

    def root_lp(milp, *, tolerance, time_limit, cupdlpx_termination):
        return O.solve_lp_relaxation(milp, tolerance=tolerance, time_limit=time_limit,
                                     cupdlpx_termination=cupdlpx_termination)

    return proj, root_lp


def pump_on_workload(name: str, time_limit: float) -> None:
    """The pump on a benchmark workload's MILP (one run of each form, at
    most `time_limit` seconds): prints one JSON line with PumpResult's time,
    pump and PDHG iterations and the root-LP, projection and rounding shares
    (src/pump.jl:412-430)."""
    from pdfp_amd.milp import is_feasible
    from pdfp_amd.pump import feasibility_pump, feasibility_pump_batch

    milp, desc = workload(name)
    out = {"workload": desc, "time_limit_s": time_limit}
    for form, run in (("fp32", lambda: feasibility_pump(milp, time_limit=time_limit,
                                                        precision="fp32")),
                      ("fp32_batch8", lambda: feasibility_pump_batch(milp, trajectories=8,
                                                                     time_limit=time_limit))):
        print(f"[bench] pump {form} on {name}", file=sys.stderr, flush=True)
        r = run()
        out[form] = {"seconds": r.time, "found": bool(r.found and is_feasible(milp, r.x)),
                     "iterations": r.iterations, "pdhg_iterations": r.pdhg_iterations,
                     "root_lp_s": r.root_lp_time, "projection_s": r.projection_time,
                     "rounding_s": r.rounding_time}
    print(json.dumps(out), flush=True)


def time_to_first_feasible(runs: int, cpu: bool, large: bool = True):
    """The full pump (root LP, projections, rounding) on the 133 × 201 binary
    set-covering instance (seed 1), on p0201 and on 2000 × 4000 and
    10,000 × 20,000 mixed MILPs (seed 1; 40k and 200k nonzeros): median PumpResult.time over `runs` runs after one warm-up run,
    as the reference times it (scripts/solve.jl:88-97); FP32 and FP64 single
    trajectories and 8 concurrent trajectories on the batched projector
    (first feasible wins, as scripts/solve.jl:121-148); and the four example
    instances' pumps together on one pack, against the same pumps one after
    another."""
    from pdfp_amd import generators as G
    from pdfp_amd.milp import is_feasible, read_mps
    from pdfp_amd.pump import feasibility_pump, feasibility_pump_batch

    out = {}
    cases = [("set_cover_133x201_seed1", G.set_cover(133, 201, 1)),
             ("p0201", read_mps(os.path.join(ROOT, "tests", "golden", "examples",
                                             "p0201.mps.gz"))),
             ("mixed_2000x4000_seed1", G.mixed(2000, 4000, 1)),
             ("mixed_10000x20000_seed1", G.mixed(10_000, 20_000, 1))]
    feasibility_pump(G.set_cover(133, 201, 7), time_limit=60.0)        # warm-up
    feasibility_pump_batch(G.set_cover(133, 201, 7), time_limit=60.0)
    # "fp32" and "fp32_batch8" finish each FP32 projection with the FP64
    # polish (the default); "fp32_unpolished" hands the FP32 point to the
    # rounder directly
    forms = {"fp32": lambda m: feasibility_pump(m, time_limit=60.0, precision="fp32"),
             "fp32_unpolished": lambda m: feasibility_pump(m, time_limit=60.0, precision="fp32",
                                                           polish=False),
             "fp64": lambda m: feasibility_pump(m, time_limit=60.0, precision="fp64"),
             "fp32_batch8": lambda m: feasibility_pump_batch(m, trajectories=8,
                                                             time_limit=60.0)}
    for name, milp in cases:
        rec = {}
        for form, run in forms.items():
            rs = [run(milp) for _ in range(runs)]
            ok = all(r.found and is_feasible(milp, r.x) for r in rs)
            ts = sorted(r.time for r in rs)
            r0 = rs[0]
            rec[form] = {"seconds": ts[len(ts) // 2], "found": ok, "iterations": r0.iterations,
                         "pdhg_iterations": r0.pdhg_iterations,
                         "root_lp_s": r0.root_lp_time, "projection_s": r0.projection_time,
                         "rounding_s": r0.rounding_time}
            if form == "fp32_batch8":
                rec[form]["winning_trajectory"] = r0.trajectory
        if cpu:
            proj, root_lp = _oracle_pump_backends()
            r = feasibility_pump(milp, time_limit=120.0, projector=proj, root_lp=root_lp,
                                 scaling="host")
            rec["cpu_oracle"] = {"seconds": r.time, "found": bool(r.found), "cores": 1,
                                 "kind": "port"}
        out[name] = rec
    out["p0201"]["reference_h100_seconds"] = 0.12   # README.md:47-48
    if large:
        # The 1M-nonzero mixed MILP of the mixed-1m workload: the polished
        # forms (median of `runs`); the unpolished FP32 pump takes 10-22 s
        # here and the oracle pump many minutes on one core, so both are left
        # out.
        milp = G.mixed(50_000, 100_000, 1)
        rec = {}
        for form in ("fp32", "fp32_batch8"):
            rs = [forms[form](milp) for _ in range(runs)]
            ts = sorted(r.time for r in rs)
            r0 = rs[0]
            rec[form] = {"seconds": ts[len(ts) // 2],
                         "found": all(r.found and is_feasible(milp, r.x) for r in rs),
                         "iterations": r0.iterations, "pdhg_iterations": r0.pdhg_iterations,
                         "root_lp_s": r0.root_lp_time, "projection_s": r0.projection_time,
                         "rounding_s": r0.rounding_time, "runs": runs}
            if form == "fp32_batch8":
                rec[form]["winning_trajectory"] = r0.trajectory
        out["mixed_50000x100000_seed1"] = rec
    # The four example instances at once: their pumps on one block-diagonal
    # pack, against the same pumps run one after another (single FP32).
    from pdfp_amd.pump import feasibility_pump_pack

    names = ("p0201", "gen-ip054", "pk1", "markshare2")
    ex = [read_mps(os.path.join(ROOT, "tests", "golden", "examples", f"{n}.mps.gz"))
          for n in names]
    feasibility_pump_pack(ex[:2], time_limit=60.0)                   # warm-up
    packs = [feasibility_pump_pack(ex, time_limit=60.0) for _ in range(runs)]
    alls = sorted(max(r.time for r in rs) for rs in packs)
    seqs = []
    for _ in range(runs):
        t = time.perf_counter()
        rs = [feasibility_pump(m, time_limit=60.0, precision="fp32") for m in ex]
        seqs.append((time.perf_counter() - t, rs))
    seqs.sort(key=lambda v: v[0])
    out["examples_4_pack"] = {
        "instances": list(names),
        "pack_all_found_seconds": alls[len(alls) // 2],
        "pack_found": all(r.found and is_feasible(m, r.x) for rs in packs
                          for m, r in zip(ex, rs)),
        "pack_instance_seconds": {n: r.time for n, r in zip(names, packs[0])},
        "sequential_fp32_seconds": seqs[len(seqs) // 2][0],
        "sequential_found": all(r.found for _, rs in seqs for r in rs),
    }
    return out


def copy_calibration(nbytes: int = 1 << 30, reps: int = 10):
    """A plain device-to-device copy of known size: the counters of its
    dispatches give the scale of FETCH_SIZE and WRITE_SIZE."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}),
          flush=True)


def bench_pack(args) -> None:
    """``--workload setcover-pack-64``: 64 set-covering instances of uneven sizes (seeds
    1..64; rows e^U[ln 2k, ln 40k], columns e^U[ln 4k, ln 80k]) in one
    block-diagonal pack.  A step is one warm-started ℓ1 projection of every
    instance to a new 0/1 target at the tolerance, all solved together; the
    value is Σ PDHG iterations over all instances / wall time.  Beside it,
    the same projections one instance at a time with the single FP32 and the
    plain FP64 projector."""
    from pdfp_amd import generators as G
    from pdfp_amd import pdfp as P
    from pdfp_amd.pack import PDFPPack

    t0 = time.perf_counter()
    milps = G.uneven_set_covers(64)
    pps = [P.projector_problem(m) for m in milps]
    mm = sum(m.A.shape[0] for m in milps)
    nn = sum(m.A.shape[1] for m in milps)
    nnz = sum(m.A.nnz for m in milps)
    desc = (f"64 set-covering instances of uneven sizes (seeds 1..64, rows e^U[ln 2k, ln 40k], "
            f"columns e^U[ln 4k, ln 80k], each column covering 3-15 rows): {mm} rows, {nn} "
            f"columns, {nnz} nnz in total, one block-diagonal pack")
    print(f"[bench] workload setcover-pack-64: {mm}x{nn}, nnz {nnz}, built in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    K = len(milps)
    p = P.SolverParameters(eps_optimal_relative=args.tol, eps_feasible_relative=args.tol,
                           iteration_limit=500_000)
    pws = [float(np.sqrt(1.0 + len(m.int_idx) / max(len(m.b), 1))) for m in milps]
    sets_w = [[G.binary_targets(m, 3000 + 100 * s + k) for k, m in enumerate(milps)]
              for s in range(args.warmup)]
    sets = [[G.binary_targets(m, 4000 + 100 * s + k) for k, m in enumerate(milps)]
            for s in range(args.steps)]

    t0 = time.perf_counter()
    pk = PDFPPack(pps)
    print(f"[bench] pack of {K} created in {time.perf_counter() - t0:.2f}s", file=sys.stderr,
          flush=True)

    def pack_steps(target_sets):
        iters, dt = 0, 0.0
        statuses = []
        for ts in target_sets:
            t = time.perf_counter()
            for k, xt in enumerate(ts):
                pk.set_target(k, xt / pps[k].D_c)
            st = pk.solve(p, pws)
            dt += time.perf_counter() - t
            iters += sum(pk.last_total_count)
            statuses.append(sorted(set(st)))
        return iters, dt, statuses

    t0 = time.perf_counter()
    pack_steps(sets_w)
    print(f"[bench] pack warmup {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    it, dt, st = pack_steps(sets)
    ms_at, ms_a = pk.time_kernels(100)
    ms_iter = pk.time_iterations(200)
    info = pk.info(0)
    print(f"[bench] fp32 pack of {K}: {it} iterations in {dt:.3f}s = {it / dt:.0f} it/s; kernels "
          f"AT {ms_at * 1e3:.1f}us A {ms_a * 1e3:.1f}us; pack iteration {ms_iter * 1e3:.1f}us; "
          f"{st}", file=sys.stderr, flush=True)
    npb = {"A": info.npb_a, "At": info.npb_at, "A_persistent_wg_per_cu": info.persist_a,
           "At_persistent_wg_per_cu": info.persist_at}
    pk.close()

    # the same projections one instance at a time
    seq = {}
    for prec in (["fp32"] if args.no_fp64 else ["fp32", "fp64"]):
        iters, secs = 0, 0.0
        t0 = time.perf_counter()
        for k, (m, pp) in enumerate(zip(milps, pps)):
            s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                            precision=prec)
            run_projections(s, pp, m, [ts[k] for ts in sets_w], args.tol)
            i, d, _ = run_projections(s, pp, m, [ts[k] for ts in sets], args.tol)
            iters += i
            secs += d
            s.close()
        seq[prec] = iters / secs
        print(f"[bench] {prec} one instance at a time: {iters} iterations in {secs:.3f}s = "
              f"{iters / secs:.0f} it/s ({time.perf_counter() - t0:.1f}s with set-up)",
              file=sys.stderr, flush=True)

    value = it / dt
    bat, ba = kernel_bytes_fp32(mm, nn, nnz)
    bat += 4 * nn  # the column's instance index
    ba += 4 * mm   # the row's instance index
    keys = ("EpiPackPrimalY<false>", "EpiPackDualZ<false>")
    dom_bytes, dom_ms, dom_name, dom_key = max(
        (bat, ms_at, "transpose pass (A^T z + x/y update)", keys[0]),
        (ba, ms_a, "row pass (A xbar + z update)", keys[1]), key=lambda t: t[1])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "PDHG iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc,
                   "step": (f"one warm-started l1 projection of each of the {K} instances to a "
                            f"new 0/1 target at relative KKT tolerance {args.tol:g}, solved "
                            f"together"),
                   "instances": K,
                   "precision": "FP32 data/iterates, FP64 reductions and polish",
                   "iterations_timed": it},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("setcover-pack-64", dom_key), "algorithmic_bytes": dom_bytes,
                     "kernel_ms": dom_ms},
        "iteration_us": ms_iter * 1e3,
        "kernel_us": {"transpose_pass": ms_at * 1e3, "row_pass": ms_a * 1e3},
        "schedule_nnz_per_block": npb,
        "fp32_one_at_a_time_value": seq["fp32"],
        "speedup_vs_fp32_one_at_a_time": value / seq["fp32"],
    }
    if "fp64" in seq:
        out["fp64_plain_one_at_a_time_value"] = seq["fp64"]
        out["speedup_vs_fp64_plain"] = value / seq["fp64"]
    print(json.dumps(out), flush=True)


def measured_record(workload: str, kernel_key: str):
    """The committed counter summary's record of a pass
    (profiles/pmc_traffic_<workload>.md, its ```json block): HBM bytes per
    launch and the kernel instantiation they were counted on; or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload.replace('-', '_')}.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if "```json" not in text:
        return None
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return json.loads(block).get(kernel_key)


def measured_traffic(workload: str, kernel_key: str):
    """HBM bytes per launch of a pass from the committed counter summary, or None."""
    rec = measured_record(workload, kernel_key)
    return None if rec is None else rec["traffic_bytes"]


def instantiation(kernel_name: str) -> tuple[str, int]:
    """(kernel template, nonzeros per block) of a CSR pass's name, e.g.
    ('csr_pass', 256) for 'csr_pass<256, 256, float, ...>'."""
    head, args = kernel_name.split("<", 1)
    return head.strip(), int(args.split(",")[1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="powerlaw-5m")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=None,
                    help="PDHG iterations of the CPU baseline (default: 400 above 2M "
                         "nonzeros, else 1000: about 10 s on one core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttff", action="store_true")
    ap.add_argument("--no-ttff-large", action="store_true",
                    help="skip the time to first feasible of the 1M-nonzero mixed MILP")
    ap.add_argument("--batch", type=int, default=8,
                    help="trajectories per step of the headline measurement (1: single)")
    ap.add_argument("--batch-schedule", default=None,
                    help="npb_a,npb_at,persist,lane_rows_a,lane_rows_at: the batched passes' "
                         "schedule instead of the autotuner's (to profile a given form)")
    ap.add_argument("--schedule", default=None,
                    help="npb_a,npb_at,persist_a,persist_at: the single-trajectory FP32 "
                         "passes' schedule instead of the autotuner's (to profile a given form)")
    ap.add_argument("--stream-launches", action="store_true",
                    help="launch every block's passes on the stream instead of replaying "
                         "captured block graphs (same kernels; for profilers that do not "
                         "trace graph replays)")
    ap.add_argument("--cooperative", type=int, default=None,
                    help="workgroups per CU of a cooperative launch running the middle "
                         "iterations of each block of the batched passes (-1: one fewer "
                         "than fit; 0: launch every pass)")
    ap.add_argument("--column-slices", choices=("on", "off", "256", "512"), default=None,
                    help="run the batch's A pass over A's 8 column slices (on) or in one "
                         "pass (off) instead of the autotuner's choice")
    ap.add_argument("--slice-min-len", type=int, default=None,
                    help="with column slices, run only the rows with more than this many "
                         "nonzeros over the slices (0: every row) instead of the "
                         "autotuner's choice")
    ap.add_argument("--batch-nt", choices=("on", "off"), default=None,
                    help="the batched passes' operand loads: non-temporal (on) or plain "
                         "(off) instead of the autotuner's choice")
    ap.add_argument("--pump", action="store_true",
                    help="only time the pump to its first feasible point on the workload's "
                         "MILP (FP32 polished, then 8 trajectories), each for at most "
                         "--pump-time-limit seconds")
    ap.add_argument("--pump-time-limit", type=float, default=120.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device 10 times (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.pump:
        pump_on_workload(args.workload, args.pump_time_limit)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the projector has no multi-GPU path in this build "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "setcover-pack-64":
        bench_pack(args)
        return

    from pdfp_amd import generators as G
    from pdfp_amd import pdfp as P

    t0 = time.perf_counter()
    milp, desc = workload(args.workload)
    P.projector_problem(milp)                      # warm-up: first use of the device
    t1 = time.perf_counter()
    pp = P.projector_problem(milp)                 # the matrix equilibrated on the device
    t2 = time.perf_counter()
    pph = P.projector_problem(milp, scaling="host")
    t3 = time.perf_counter()
    if not (np.array_equal(pp.A.data, pph.A.data) and np.array_equal(pp.D_r, pph.D_r)
            and np.array_equal(pp.D_c, pph.D_c)):
        raise SystemExit("device and host scaling differ")
    scaling_s = {"device": t2 - t1, "host_numpy": t3 - t2}
    del pph
    mm, nn = milp.A.shape
    nnz = milp.A.nnz
    print(f"[bench] workload {args.workload}: {mm}x{nn}, nnz {nnz}, built in "
          f"{time.perf_counter() - t0:.1f}s; projector scaling {scaling_s['device']:.3f}s on the "
          f"device, {scaling_s['host_numpy']:.3f}s on the host", file=sys.stderr, flush=True)
    targets_w = [G.binary_targets(milp, 1000 + k) for k in range(args.warmup)]
    targets = [G.binary_targets(milp, 2000 + k) for k in range(args.steps)]

    results = {}
    for prec in (["fp32"] if args.no_fp64 else ["fp32", "fp64"]):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision=prec)
        if args.stream_launches:
            s.set_block_graph(False)
        if args.schedule and prec == "fp32":
            s.set_schedule(*[int(v) for v in args.schedule.split(",")])
        s.set_x(np.clip(np.zeros(nn), milp.l, milp.u) / pp.D_c)
        t0 = time.perf_counter()
        run_projections(s, pp, milp, targets_w, args.tol)  # warmup (includes ‖A‖₂)
        print(f"[bench] {prec} warmup {time.perf_counter() - t0:.2f}s", file=sys.stderr,
              flush=True)
        it, dt, st = run_projections(s, pp, milp, targets, args.tol)
        quality = run_projections.quality
        ms_at, ms_a = s.time_kernels(200)
        ms_iter = s.time_iterations(400)
        info = s.info()
        spmv_ms = ((s.time_spmv(100), s.time_spmv(100, transpose=True)) if prec == "fp32"
                   else None)
        results[prec] = dict(iters=it, seconds=dt, statuses=st, ms_at=ms_at, ms_a=ms_a,
                             ms_iter=ms_iter, info=info, quality=quality, spmv_ms=spmv_ms,
                             npb={"A": info.npb_a, "At": info.npb_at,
                                  "A_persistent_wg_per_cu": info.persist_a,
                                  "At_persistent_wg_per_cu": info.persist_at})
        print(f"[bench] {prec}: {it} iterations in {dt:.3f}s = {it / dt:.0f} it/s; "
              f"kernels AT {ms_at * 1e3:.1f}us A {ms_a * 1e3:.1f}us; iteration "
              f"{ms_iter * 1e3:.1f}us; {st}", file=sys.stderr, flush=True)
        s.close()

    K = args.batch
    if K > 1:
        from pdfp_amd.batch import PDFPBatch

        bt = PDFPBatch(K, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
        if args.batch_schedule:
            bt.set_schedule(*[int(v) for v in args.batch_schedule.split(",")])
        if args.stream_launches:
            bt.set_block_graph(False)
        if args.cooperative is not None:
            bt.set_cooperative(args.cooperative)
        if args.column_slices is not None:
            bt.set_column_slices({"on": 1, "off": 0, "256": 256,
                                  "512": 512}[args.column_slices])
        if args.slice_min_len is not None:
            bt.set_slice_min_len(args.slice_min_len)
        if args.batch_nt is not None:
            bt.set_nontemporal(args.batch_nt == "on")
        for k in range(K):
            bt.set_x(k, np.clip(np.zeros(nn), milp.l, milp.u) / pp.D_c)
        sets_w = [[G.binary_targets(milp, 3000 + 100 * s + k) for k in range(K)]
                  for s in range(args.warmup)]
        sets = [[G.binary_targets(milp, 4000 + 100 * s + k) for k in range(K)]
                for s in range(args.steps)]
        t0 = time.perf_counter()
        run_batch(bt, pp, milp, sets_w, args.tol)
        print(f"[bench] batch{K} warmup {time.perf_counter() - t0:.2f}s", file=sys.stderr,
              flush=True)
        it, dt, st = run_batch(bt, pp, milp, sets, args.tol)
        ms_at, ms_a = bt.time_kernels(100)
        ms_iter = bt.time_iterations(200)
        bi = bt.info(0)
        results["batch"] = dict(iters=it, seconds=dt, statuses=st, ms_at=ms_at, ms_a=ms_a,
                                ms_iter=ms_iter,
                                npb={"A": bi.npb_a, "At": bi.npb_at,
                                     "A_persistent_wg_per_cu": bi.persist_a,
                                     "At_persistent_wg_per_cu": bi.persist_at,
                                     "A_lane_rows": bi.lane_rows_a,
                                     "At_lane_rows": bi.lane_rows_at,
                                     "A_column_slices": bi.column_slices,
                                     "A_slice_min_len": bi.slice_min_len,
                                     "nontemporal": bi.nontemporal})
        print(f"[bench] fp32 batch of {K}: {it} iterations in {dt:.3f}s = {it / dt:.0f} it/s; "
              f"kernels AT {ms_at * 1e3:.1f}us A {ms_a * 1e3:.1f}us; batched iteration "
              f"{ms_iter * 1e3:.1f}us", file=sys.stderr, flush=True)
        bt.close()
        head = results["batch"]
        bat, ba = kernel_bytes_batch(mm, nn, nnz, K)
        if bi.column_slices:
            # the A pass runs over column slices: its own layout's bytes (the
            # one-pass figure stays beside them in the line)
            head["row_pass_bytes_one_pass"] = ba
            T, n_long = slice_rows(pp.A, bi.slice_min_len)
            lens = np.diff(pp.A.tocsr().indptr)
            ba = kernel_bytes_sliced_a(mm, nn, nnz, T, K, min_len=bi.slice_min_len,
                                       nnz_long=int(lens[lens > bi.slice_min_len].sum()),
                                       n_long=n_long)
        steps_desc = (f"{K} warm-started l1 projections to new integer targets at relative KKT "
                      f"tolerance {args.tol:g}, served by the {K} trajectory slots of one batch "
                      f"sharing A (a slot that finishes takes the next target)")
    else:
        head = results["fp32"]
        bat, ba = kernel_bytes_fp32(mm, nn, nnz)
        steps_desc = (f"one warm-started l1 projection to a new 0/1 target at relative KKT "
                      f"tolerance {args.tol:g}")
    r32 = results["fp32"]
    value = head["iters"] / head["seconds"]
    width = 8 if K <= 8 else 16
    nt = "NT" if K > 1 and head["npb"].get("nontemporal") else ""
    keys = ((f"EpiBatchPrimalY{nt}<{width}, false>", f"EpiBatchDualZ{nt}<{width}, false>")
            if K > 1 else ("EpiPrimalY<float, false>", "EpiDualZ<float, false>"))
    if K > 1 and head["npb"].get("A_column_slices"):
        keys = (keys[0], f"EpiSlicePartialLanes<{width}>" if head["npb"]["A_slice_min_len"]
                else f"EpiSlicePartial<{width}>")
    dom_bytes, dom_ms, dom_name, dom_key = max(
        (bat, head["ms_at"], "transpose pass (A^T z + x/y update)", keys[0]),
        (ba, head["ms_a"], "row pass (A xbar + z update)", keys[1]), key=lambda t: t[1])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    # The timed launches' instantiations (kernel template, nonzeros per
    # block), to check that the committed counts were taken on the same ones.
    sched = head["npb"]
    if K > 1:
        def grid(p):
            return "csr_pass_persistent" if p else "csr_pass"
        timed = ([(grid(sched["At_persistent_wg_per_cu"]), sched["At"])] if dom_key == keys[0]
                 else [(grid(sched["A_persistent_wg_per_cu"]), sched["A"])])
    else:
        timed = ([("csr_pass" if not sched["At_persistent_wg_per_cu"] else
                   "csr_pass_persistent", sched["At"])] if dom_key == keys[0] else
                 [("csr_pass" if not sched["A_persistent_wg_per_cu"] else
                   "csr_pass_persistent", sched["A"])])
    recs = [measured_record(args.workload, dom_key)]
    if dom_key.startswith("EpiSlicePartial"):
        # the sliced A pass is two launches: the slices' pass, then the combine
        # (every row over the slices) or the short rows' pass
        second = ((f"k_batch_slice_dualz_nt<{width}, false>" if nt else
                   f"k_batch_slice_dualz<{width}, false>") if not sched["A_slice_min_len"]
                  else f"EpiBatchDualZLong{nt}<{width}, false>")
        recs.append(measured_record(args.workload, second))
        timed = [("csr_pass_xcd_segments", sched["A_column_slices"])]
        if sched["A_slice_min_len"]:
            timed.append((grid(sched["A_persistent_wg_per_cu"]), sched["A"]))
    traffic = (None if any(r is None for r in recs)
               else sum(r["traffic_bytes"] for r in recs))
    counted_on = [r["kernel_fetch_run"] for r in recs if r is not None]
    counted = [instantiation(k) for k in counted_on if k.startswith("csr_pass")]
    same_kernels = counted == timed
    # The random-gather floor of each pass: nnz gathers of one record per
    # column (Aᵀz gathers z: m records; A x̄ gathers x̄: n records).
    from pdfp_amd._lib import gather_ceiling_us

    rec = (32 if K <= 8 else 64) if K > 1 else 4
    ceil = {"transpose_pass": gather_ceiling_us(max(mm, 1) * rec, rec, nnz),
            "row_pass": gather_ceiling_us(max(nn, 1) * rec, rec, nnz)}
    dom_ceil = ceil["transpose_pass" if dom_key == keys[0] else "row_pass"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "PDHG iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "step": steps_desc, "trajectories": K,
                   "precision": "FP32 data/iterates, FP64 reductions and polish",
                   "iterations_timed": head["iters"]},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": dom_bytes,
                     "kernel_ms": dom_ms, "gather_ceiling_us": dom_ceil,
                     "frac_of_gather_ceiling": dom_ceil / (dom_ms * 1e3),
                     # the counted bytes (line fills of the gathers included) per second
                     "traffic_gbs": (None if traffic is None
                                     else traffic / (dom_ms * 1e-3) / 1e9),
                     "traffic_frac": (None if traffic is None
                                      else traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
                     # the kernels the committed counts come from, and whether
                     # they are the timed launches' instantiations (the
                     # autotuner may pick another block size on another run)
                     "traffic_counted_on": counted_on,
                     "traffic_same_instantiation": same_kernels},
        "iteration_us": head["ms_iter"] * 1e3,
        "kernel_us": {"transpose_pass": head["ms_at"] * 1e3, "row_pass": head["ms_a"] * 1e3},
        "gather_ceiling_us": ceil,
        "fp32_single_value": r32["iters"] / r32["seconds"],
        "fp32_single_iteration_us": r32["ms_iter"] * 1e3,
        "fp32_single_kernel_us": {"transpose_pass": r32["ms_at"] * 1e3,
                                  "row_pass": r32["ms_a"] * 1e3},
        "schedule_nnz_per_block": r32["npb"],
        "scaling_seconds": scaling_s,
        "spmv": spmv_figures(mm, nn, nnz, r32["spmv_ms"]),
    }
    if K > 1:
        out["batch_schedule_nnz_per_block"] = head["npb"]
        if "row_pass_bytes_one_pass" in head:
            out["row_pass_bytes"] = {"column_slices": ba,
                                     "one_pass": head["row_pass_bytes_one_pass"]}
    if "fp64" in results:
        r64 = results["fp64"]
        v64 = r64["iters"] / r64["seconds"]
        out["fp64_plain_value"] = v64
        out["speedup_vs_fp64_plain"] = value / v64
        out["speedup_fp32_single_vs_fp64_plain"] = out["fp32_single_value"] / v64
        out["fp64_iteration_us"] = r64["ms_iter"] * 1e3
        # each FP32 projection against the FP64 form's projection of the same
        # target from the same start (the FP64 form follows the reference's
        # arithmetic and matches the C oracle's iterations)
        out["fp32_vs_fp64_per_projection"] = [
            {"l1_fp32": a["l1"], "l1_fp64": b["l1"],
             "l1_rel_diff": abs(a["l1"] - b["l1"]) / max(1.0, abs(b["l1"])),
             "fp32_rel_primal": a["rel_primal"], "fp32_rel_dual": a["rel_dual"],
             "fp32_rel_gap": a["rel_gap"]}
            for a, b in zip(r32["quality"], r64["quality"])]
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu_it = args.cpu_iterations or (400 if nnz > 2_000_000 else 1000)
        out["cpu_baseline"] = cpu_baseline(pp, milp, targets[0], cpu_it)
        # the host's share of cores on the box (at most 16 threads)
        thr = max(1, min(16, len(os.sched_getaffinity(0))))
        if thr > 1:
            print(f"[bench] cpu baseline, {thr} threads", file=sys.stderr, flush=True)
            out["cpu_baseline_threaded"] = cpu_baseline(pp, milp, targets[0], 2 * cpu_it, thr)
    if not args.no_ttff:
        print("[bench] time to first feasible", file=sys.stderr, flush=True)
        out["time_to_first_feasible"] = time_to_first_feasible(3, cpu=not args.no_cpu,
                                                               large=not args.no_ttff_large)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
