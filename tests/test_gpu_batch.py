"""GPU tests of the shared-matrix trajectory batch (pdfp_batch_*).

Each trajectory must behave as an independent FP32 projection: its KKT
residuals, re-evaluated in FP64 by the oracle, meet the tolerance (×10 for
FP32 rounding), its ℓ1 distance matches the single FP32 projector's to
1e-4 relative, and trajectories that are not active are left unchanged."""
import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd import pdfp as P
from pdfp_amd.batch import PDFPBatch
from pdfp_amd.milp import read_mps
from pdfp_amd.pdfp import SolverParameters, projector_problem

from .conftest import example_path

pytestmark = pytest.mark.gpu


def _batch(pp, K):
    return PDFPBatch(K, pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)


def _oracle_kkt(pp, xt_scaled, x, y, z, ds):
    o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    O.set_target(o, xt_scaled)
    o.initialize(None, 1.0)
    return o.residual_of(x, y, z, ds)


@pytest.mark.parametrize("K", [8, 3, 16, 11])
def test_batch_trajectories_are_independent_projections(K):
    print("batch", K, flush=True)
    m = G.mixed(400, 800, 31)
    pp = projector_problem(m)
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    bt = _batch(pp, K)
    targets = [G.binary_targets(m, 500 + k) / pp.D_c for k in range(K)]
    for k in range(K):
        bt.set_target(k, targets[k])
    terms = bt.solve(p, np.linspace(0.8, 1.5, K))
    assert terms == ["OPTIMAL"] * K
    for k in range(K):
        x = bt.pdhg_x(k)
        r = _oracle_kkt(pp, targets[k], x, bt.vector(k, "pdhg_y"), bt.vector(k, "pdhg_z"),
                        bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, float(np.linspace(0.8, 1.5, K)[k])) == "OPTIMAL"
        l1b, l1s = bt.info(k).l1_distance, s.info().l1_distance
        assert abs(l1b - l1s) <= 1e-4 * max(1.0, l1s), (k, l1b, l1s)


def test_batch_inactive_trajectory_is_untouched_and_warm_starts():
    m = G.set_cover(133, 201, 4)
    pp = projector_problem(m)
    bt = _batch(pp, 4)
    p = SolverParameters(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
    for k in range(4):
        bt.set_target(k, np.zeros(201))
    assert bt.solve(p, 1.0) == ["OPTIMAL"] * 4
    before = [bt.pdhg_x(k) for k in range(4)]
    bt.set_target(1, np.ones(201) / pp.D_c)        # feasible target: distance 0
    terms = bt.solve(p, 1.0, active=[0, 1, 0, 0])
    assert terms[1] == "OPTIMAL"
    for k in (0, 2, 3):
        assert np.array_equal(bt.pdhg_x(k), before[k])
    assert bt.info(1).l1_distance <= 1e-5
    # the same targets in all slots give the same projection
    l1 = [bt.info(k).l1_distance for k in (0, 2, 3)]
    assert max(l1) - min(l1) <= 1e-6 * max(1.0, max(l1))


def _with_dense_column(m, value=0.1):
    """m plus one continuous column in every row: a row of Aᵀ longer than
    any block, so both batched passes take their long-row (vector) path."""
    import scipy.sparse as sp

    from pdfp_amd.milp import MILP

    rows = m.A.shape[0]
    col = sp.csc_matrix(np.full((rows, 1), value))
    A = sp.hstack([m.A, col], format="csc")
    return MILP(c=np.append(m.c, 0.0), c0=m.c0, A=A, b=m.b, row_sense=m.row_sense,
                l=np.append(m.l, 0.0), u=np.append(m.u, 1.0), int_idx=m.int_idx,
                name=m.name + "_dense_col")


@pytest.mark.parametrize("K", [5, 13])
def test_batch_long_rows_and_columns(K):
    """Power-law rows (up to 3000 nonzeros) and a dense column: the
    long-row path of both passes, with 5 trajectories (width 8) and 13
    (width 16)."""
    print("batch long rows", K, flush=True)
    m = _with_dense_column(G.mixed(2000, 5000, 3, power_law=True, nnz=60_000, max_len=3000))
    pp = projector_problem(m)
    assert np.diff(pp.A.tocsr().indptr).max() > 2048 and np.diff(pp.A.indptr).max() >= 2000
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    bt = _batch(pp, K)
    targets = [G.binary_targets(m, 700 + k) / pp.D_c for k in range(K)]
    for k in range(K):
        bt.set_target(k, targets[k])
    assert bt.solve(p, 1.0) == ["OPTIMAL"] * K
    for k in range(K):
        r = _oracle_kkt(pp, targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"),
                        bt.vector(k, "pdhg_z"), bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        l1b, l1s = bt.info(k).l1_distance, s.info().l1_distance
        assert abs(l1b - l1s) <= 1e-4 * max(1.0, l1s), (k, l1b, l1s)


def test_batch_continuous_begin_and_run():
    """pdfp_batch_begin / pdfp_batch_run: trajectories started at different
    times, each with its own tolerance, end as their own solves would, and a
    trajectory that ended can start again while another still runs."""
    print("batch continuous", flush=True)
    m = G.mixed(400, 800, 31)
    pp = projector_problem(m)
    bt = _batch(pp, 3)
    tg = [G.binary_targets(m, 900 + k) / pp.D_c for k in range(3)]
    loose = SolverParameters(eps_optimal_relative=1e-2, eps_feasible_relative=1e-3)
    tight = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                             iteration_limit=200_000)
    bt.set_target(0, tg[0])
    bt.begin(0, tight, 1.0)
    bt.set_target(1, tg[1])
    bt.begin(1, loose, 1.0)
    terms, left = bt.run(until_any=True)
    assert terms[1] == "OPTIMAL" and terms[0] == "UNSPECIFIED" and left == 1
    n1 = bt.total_count(1)
    bt.set_target(2, tg[2])                       # joins while 0 still runs
    bt.begin(2, tight, 1.0)
    while left:
        terms, left = bt.run(until_any=True)
    assert terms[0] == terms[2] == "OPTIMAL" and bt.total_count(1) == n1
    for k in (0, 2):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, tg[k])
        assert P.pdfp_solve(s, tight, 1.0) == "OPTIMAL"
        assert abs(bt.info(k).l1_distance - s.info().l1_distance) <= \
            1e-4 * max(1.0, s.info().l1_distance)
    bad = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                           termination_evaluation_frequency=100)
    bt.set_target(0, tg[0])
    bt.begin(0, tight, 1.0)
    bt.set_target(1, tg[1])
    bt.begin(1, bad, 1.0)
    from pdfp_amd._lib import LibraryError

    with pytest.raises(LibraryError, match="share"):
        bt.run()


def test_batch_one_block_runs_match_until_any():
    """pdfp_batch_run with until_any = 2 returns after every block: stepped
    block by block, two trajectories reach the same terminations, iteration
    counts and points, bit for bit, as when each call runs until one ends."""
    print("batch one-block runs", flush=True)
    m = G.mixed(400, 800, 32)
    pp = projector_problem(m)
    tg = [G.binary_targets(m, 950 + k) / pp.D_c for k in range(2)]
    p = [SolverParameters(eps_optimal_relative=1e-3, eps_feasible_relative=1e-4),
         SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                          iteration_limit=200_000)]
    out = []
    for one_block in (False, True):
        bt = _batch(pp, 2)
        bt.set_kernel_path("general")
        bt.set_schedule(256, 256, 0, 32, 32)
        for k in range(2):
            bt.set_target(k, tg[k])
            bt.begin(k, p[k], 1.0)
        left, calls, prev = 2, 0, [0, 0]
        while left:
            terms, left = bt.run(until_any=True, one_block=one_block)
            calls += 1
            if one_block:   # each call is one block of 200 for every running one
                counts = bt.last_total_count
                assert all(counts[k] - prev[k] in (0, 200) for k in range(2))
                assert max(counts[k] - prev[k] for k in range(2)) == 200
                prev = counts
        out.append((terms, [bt.total_count(k) for k in range(2)],
                    [bt.pdhg_x(k) for k in range(2)], calls))
        bt.close()
    (t0, c0, x0, n0), (t1, c1, x1, n1) = out
    assert t0 == t1 == ["OPTIMAL", "OPTIMAL"] and c0 == c1
    assert all(np.array_equal(a, b) for a, b in zip(x0, x1))
    assert n1 == max(c1) // 200 and n0 < n1
    bt = _batch(pp, 2)
    assert bt._lib.pdfp_batch_run(bt._ctx, 3, None, None, None) != 0   # no such mode
    assert bt._lib.pdfp_batch_run(bt._ctx, -1, None, None, None) != 0
    bt.close()


@pytest.mark.parametrize("K", [4, 10])
def test_batch_schedules_agree(K):
    """Every schedule form of the batched passes (nonzeros per workgroup,
    persistent grid, lane-per-(row, trajectory) reduction of short-row
    blocks, up to every stream-mode block) gives the same projections:
    oracle KKT within 10×tol and the single FP32 projector's ℓ1 distance to
    1e-4, on power-law rows with a dense column; batch widths 8 and 16."""
    m = _with_dense_column(G.mixed(2000, 5000, 5, power_law=True, nnz=60_000, max_len=3000))
    pp = projector_problem(m)
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    targets = [G.binary_targets(m, 40 + k) / pp.D_c for k in range(K)]
    ref = []
    for k in range(K):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        ref.append(s.info().l1_distance)
    bt = _batch(pp, K)
    for sched in [(256, 256, 0, 0, 0), (512, 512, 0, 32, 32), (256, 512, 4, 128, 128),
                  (512, 256, 4, 1 << 30, 1 << 30), (512, 512, 0, 1 << 30, 0)]:
        print("batch schedule", sched, flush=True)
        bt.set_schedule(*sched)
        info = bt.info(0)
        assert (info.npb_a, info.npb_at, info.lane_rows_a, info.lane_rows_at) == \
            (sched[0], sched[1], sched[3], sched[4])
        for k in range(K):
            bt.set_target(k, targets[k])
        assert bt.solve(p, 1.0) == ["OPTIMAL"] * K, sched
        for k in range(K):
            r = _oracle_kkt(pp, targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"),
                            bt.vector(k, "pdhg_z"), bt.vector(k, "dual_slack"))
            assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (sched, k, r)
            l1b = bt.info(k).l1_distance
            assert abs(l1b - ref[k]) <= 1e-4 * max(1.0, ref[k]), (sched, k, l1b, ref[k])


@pytest.mark.parametrize("min_len", [0, 64, 100_000])
@pytest.mark.parametrize("K", [8, 16, 5])
def test_batch_column_slices_agree(K, min_len):
    """The A pass over A's 8 column slices (one per XCD, FP32 partial sums
    added in FP64) gives the same projections as the one-pass form: oracle
    KKT within 10×tol and the single FP32 projector's ℓ1 distance to 1e-4,
    on power-law rows with a dense column; the replayed block graphs equal
    the stream launches bit for bit.  min_len 64: only the rows longer than
    64 nonzeros run over the slices, the others in one pass; 100,000: no row
    is that long, so the slices' pass has nothing to do."""
    print("batch column slices", K, min_len, flush=True)
    m = _with_dense_column(G.mixed(2000, 5000, 5, power_law=True, nnz=60_000, max_len=3000))
    pp = projector_problem(m)
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    targets = [G.binary_targets(m, 60 + k) / pp.D_c for k in range(K)]
    ref = []
    for k in range(K):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, targets[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        ref.append(s.info().l1_distance)
    runs = []
    for graph in (True, False):
        bt = _batch(pp, K)
        bt.set_kernel_path("general")
        bt.set_schedule(512, 256, 0, 32, 32)
        bt.set_column_slices(True)
        bt.set_slice_min_len(min_len)
        bt.set_block_graph(graph)
        for k in range(K):
            bt.set_target(k, targets[k])
        assert bt.solve(p, 1.0) == ["OPTIMAL"] * K
        assert bt.info(0).column_slices == 512
        assert bt.info(0).slice_min_len == min_len
        run = []
        for k in range(K):
            x, z = bt.pdhg_x(k), bt.vector(k, "pdhg_z")
            r = _oracle_kkt(pp, targets[k], x, bt.vector(k, "pdhg_y"), z,
                            bt.vector(k, "dual_slack"))
            assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
            l1b = bt.info(k).l1_distance
            assert abs(l1b - ref[k]) <= 1e-4 * max(1.0, ref[k]), (k, l1b, ref[k])
            run.append((bt.info(k).total_count, x, z))
        bt.set_column_slices(False)
        assert bt.info(0).column_slices == 0
        bt.close()
        runs.append(run)
    for a, b in zip(*runs):
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("slices", [0, 64])
@pytest.mark.parametrize("K", [8, 16])
def test_batch_nontemporal_loads_change_nothing(K, slices):
    """The passes' operands by non-temporal loads give the same trajectories,
    bit for bit, as by plain loads (one pass, and the long rows over the
    column slices), and pdfp_info reports the choice."""
    print("batch nontemporal", K, slices, flush=True)
    m = _with_dense_column(G.mixed(2000, 5000, 5, power_law=True, nnz=60_000, max_len=3000))
    pp = projector_problem(m)
    p = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                         iteration_limit=200_000)
    targets = [G.binary_targets(m, 70 + k) / pp.D_c for k in range(K)]
    runs = []
    for nt in (False, True):
        bt = _batch(pp, K)
        bt.set_kernel_path("general")
        bt.set_schedule(512, 256, 0, 32, 32)
        if slices:
            bt.set_column_slices(True)
            bt.set_slice_min_len(slices)
        bt.set_nontemporal(nt)
        for k in range(K):
            bt.set_target(k, targets[k])
        assert bt.solve(p, 1.0) == ["OPTIMAL"] * K
        assert bt.info(0).nontemporal == int(nt)
        runs.append([(bt.info(k).total_count, bt.pdhg_x(k), bt.vector(k, "pdhg_z"))
                     for k in range(K)])
        bt.close()
    for a, b in zip(*runs):
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("K", [8, 16])
def test_batch_block_graph_replays_equal_stream_launches(K):
    """Blocks replayed as captured graphs give the same trajectories, bit for
    bit, as the same launches on the stream: trajectories that start, end,
    restart and start again at different blocks (so the device masks and
    the step tables change under the replayed graph), each run on a fresh
    batch with the same schedule."""
    print("batch block graph", K, flush=True)
    m = G.mixed(2000, 4000, 1)
    pp = projector_problem(m)
    tg = [G.binary_targets(m, 700 + k) / pp.D_c for k in range(K + 4)]
    loose = SolverParameters(eps_optimal_relative=1e-3, eps_feasible_relative=1e-4)
    tight = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                             iteration_limit=200_000)
    runs = []
    for graph in (False, True, True):
        bt = _batch(pp, K)
        bt.set_kernel_path("general")
        bt.set_schedule(256, 256, 0, 32, 32)
        bt.set_block_graph(graph)
        queue = list(range(K + 4))
        ended = []
        running = [True] * K
        for k in range(K):
            bt.set_target(k, tg[queue.pop(0)])
            bt.begin(k, tight if k % 2 else loose, 0.6 + 0.1 * k)
        while any(running):
            terms, _ = bt.run(until_any=True)
            for k in range(K):
                if not running[k] or terms[k] == "UNSPECIFIED":
                    continue
                assert terms[k] == "OPTIMAL"
                info = bt.info(k)
                assert info.block_graph == int(graph)
                ended.append((k, info.total_count, info.n_restarts, info.l1_distance,
                              bt.vector(k, "pdhg_x"), bt.vector(k, "pdhg_z")))
                running[k] = False
                if queue:
                    bt.set_target(k, tg[queue.pop(0)])
                    bt.begin(k, loose, 1.0)
                    running[k] = True
        bt.close()
        runs.append(ended)
    assert len(runs[0]) == K + 4
    for run in runs[1:]:
        assert len(run) == len(runs[0])
        for a, b in zip(runs[0], run):
            assert a[:4] == b[:4]
            assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


@pytest.mark.parametrize("K,sched", [(8, (256, 256, 0, 32, 32)), (16, (512, 256, 0, 0, 0)),
                                     (5, (512, 512, 4, 128, 128))])
def test_batch_cooperative_blocks_equal_stream_launches(K, sched):
    """Blocks whose middle iterations run in one cooperative launch (grid
    barriers between the passes) give the same trajectories, bit for bit, as
    the same passes launched one by one: trajectories that start, end,
    restart and start again at different blocks, with evaluation
    frequencies of 1, 2, 3 and 40 (no cooperative launch, none, one
    iteration in it, the usual case)."""
    print("batch cooperative", K, sched, flush=True)
    m = G.mixed(2000, 4000, 1)
    pp = projector_problem(m)
    tg = [G.binary_targets(m, 900 + k) / pp.D_c for k in range(K + 3)]
    for freq in (40, 3, 2, 1):
        loose = SolverParameters(eps_optimal_relative=1e-3, eps_feasible_relative=1e-4,
                                 termination_evaluation_frequency=freq)
        tight = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                                 iteration_limit=200_000 if freq > 3 else 4_000,
                                 termination_evaluation_frequency=freq)
        runs = []
        for coop in (False, True):
            print("  freq", freq, "cooperative", coop, flush=True)
            bt = _batch(pp, K)
            bt.set_kernel_path("general")
            bt.set_schedule(*sched)
            bt.set_cooperative(coop)
            queue = list(range(K + 3))
            ended = []
            running = [True] * K
            for k in range(K):
                bt.set_target(k, tg[queue.pop(0)])
                bt.begin(k, tight if k % 2 else loose, 0.6 + 0.1 * k)
            while any(running):
                terms, _ = bt.run(until_any=True)
                for k in range(K):
                    if not running[k] or terms[k] == "UNSPECIFIED":
                        continue
                    assert terms[k] in ("OPTIMAL", "ITERATION_LIMIT")
                    info = bt.info(k)
                    ended.append((k, terms[k], info.total_count, info.n_restarts,
                                  info.l1_distance, bt.vector(k, "pdhg_x"),
                                  bt.vector(k, "pdhg_z")))
                    running[k] = False
                    if queue:
                        bt.set_target(k, tg[queue.pop(0)])
                        bt.begin(k, loose, 1.0)
                        running[k] = True
            if coop:
                assert bt.time_iterations(50) > 0.0
            bt.close()
            runs.append(ended)
        assert len(runs[0]) == K + 3 == len(runs[1])
        for a, b in zip(runs[0], runs[1]):
            assert a[:5] == b[:5], (freq, a[:5], b[:5])
            assert np.array_equal(a[5], b[5]) and np.array_equal(a[6], b[6]), (freq, a[0])


@pytest.mark.parametrize("K", [8, 16])
def test_batch_small_instance_trajectories_equal_single_projector_bitwise(K):
    """When the instance fits in LDS, the batch runs one workgroup per running
    trajectory on its view of the interleaved vectors, each iterating as the
    single FP32 projector's one-workgroup path: every trajectory's
    projections (a cold one, then a warm-started one begun at another
    block) equal the single projector's bit for bit, and the batched passes
    (path forced) reach the same ℓ1 distances to the tolerance."""
    print("batch small", K, flush=True)
    m = read_mps(example_path("p0201"))
    pp = projector_problem(m)
    tol = 1e-6
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=100_000)
    pw = [0.7 + 0.1 * k for k in range(K)]
    t1 = [G.binary_targets(m, 60 + k) / pp.D_c for k in range(K)]
    t2 = [G.binary_targets(m, 80 + k) / pp.D_c for k in range(K)]
    ref = []
    for k in range(K):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        seq = []
        for t in (t1[k], t2[k]):
            P.set_target(s, t)
            assert P.pdfp_solve(s, p, pw[k]) == "OPTIMAL"
            assert s.info().small_path == 1
            seq.append((s.total_count, s.info().l1_distance, s.pdhg_x.copy()))
        s.close()
        ref.append(seq)
    bt = _batch(pp, K)
    for k in range(K):
        bt.set_target(k, t1[k])
        bt.begin(k, p, pw[k])
    running, second = [True] * K, [False] * K
    got = [[] for _ in range(K)]
    while any(running):
        terms, _ = bt.run(until_any=True)
        for k in range(K):
            if not running[k] or terms[k] == "UNSPECIFIED":
                continue
            assert terms[k] == "OPTIMAL" and bt.info(k).small_path == 1
            got[k].append((bt.total_count(k), bt.info(k).l1_distance, bt.pdhg_x(k)))
            running[k] = False
            if not second[k]:
                second[k] = running[k] = True
                bt.set_target(k, t2[k])
                bt.begin(k, p, pw[k])
    for k in range(K):
        for a, b in zip(ref[k], got[k]):
            assert a[0] == b[0] and a[1] == b[1], (k, a[:2], b[:2])
            assert np.array_equal(a[2], b[2]), k
    bt.set_kernel_path("general")
    for k in range(K):
        bt.set_target(k, t1[k])
    assert bt.solve(p, np.asarray(pw)) == ["OPTIMAL"] * K
    for k in range(K):
        assert bt.info(k).small_path == 0
        l1g, l1s = bt.info(k).l1_distance, ref[k][0][1]
        assert abs(l1g - l1s) <= 1e-4 * max(1.0, l1s), (k, l1g, l1s)


@pytest.mark.parametrize("path", ["auto", "general"])
def test_batch_on_instance_with_empty_row_and_column(path):
    """Two trajectories on an instance with an empty row, an equality row
    and an empty column with an infinite upper bound, on both kernel paths:
    each ends as the single FP32 projector does on the same target."""
    import scipy.sparse as sp

    print("batch edges", path, flush=True)
    A = sp.csc_matrix(np.array([[1.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0],
                                [0.0, 1.0, 1.0, 0.0]]))
    m = G.MILP(c=np.zeros(4), c0=0.0, A=A, b=np.array([1.0, 1.0, 1.0]),
               row_sense=np.array(["le", "le", "eq"]), l=np.zeros(4),
               u=np.array([1.0, 1.0, 1.0, np.inf]), int_idx=np.array([0, 1, 2]))
    pp = projector_problem(m)
    p = SolverParameters(eps_optimal_relative=1e-5, eps_feasible_relative=1e-5,
                         iteration_limit=200_000)
    tg = [np.array([1.0, 1.0, 1.0, 5.0]) / pp.D_c, np.array([0.0, 1.0, 1.0, 0.0]) / pp.D_c]
    bt = _batch(pp, 2)
    bt.set_kernel_path(path)
    for k in range(2):
        bt.set_target(k, tg[k])
    assert bt.solve(p, np.ones(2)) == ["OPTIMAL"] * 2
    for k in range(2):
        s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                        precision="fp32")
        P.set_target(s, tg[k])
        assert P.pdfp_solve(s, p, 1.0) == "OPTIMAL"
        l1b, l1s = bt.info(k).l1_distance, s.info().l1_distance
        assert abs(l1b - l1s) <= 1e-4 * max(1.0, l1s), (k, l1b, l1s)
        s.close()
    bt.close()


def test_batch_general_integer_equality_heavy():
    """Eight trajectories on a general-integer, equality-heavy instance (40 %
    equality rows, integer columns in [0, U_j] with U_j up to 20, integer
    coefficients): each trajectory's point meets the tolerance under the
    oracle's FP64 residuals, and its ℓ1 distance is within the gap tolerance
    of the oracle's own FP64 projection of the same target."""
    print("batch general integer", flush=True)
    m = G.general_integer(2000, 4000, 5)
    assert (m.row_sense == "eq").mean() > 0.3 and m.u.max() > 1
    pp = projector_problem(m)
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    pw = float(np.sqrt(1.0 + len(m.int_idx) / len(m.b)))
    bt = _batch(pp, 8)
    targets = [G.binary_targets(m, 700 + k) / pp.D_c for k in range(8)]
    for k in range(8):
        bt.set_target(k, targets[k])
    assert bt.solve(p, pw) == ["OPTIMAL"] * 8
    ii = np.asarray(pp.is_int, dtype=bool)
    for k in range(8):
        r = _oracle_kkt(pp, targets[k], bt.pdhg_x(k), bt.vector(k, "pdhg_y"),
                        bt.vector(k, "pdhg_z"), bt.vector(k, "dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
        if k < 3:
            o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
            O.set_target(o, targets[k])
            assert O.pdfp_solve(o, p, pw) == "OPTIMAL"
            l1o = o.info().l1_distance
            den = 1.0 + float(np.sum(np.asarray(pp.w)[ii] * np.abs(targets[k][ii])))
            assert abs(r[3] - l1o) <= 2 * 10 * tol * den, (k, r[3], l1o)


@pytest.mark.parametrize("name,maker,path", [
    ("p0201", lambda: read_mps(example_path("p0201")), "auto"),
    ("mixed_2000x4000", lambda: G.mixed(2000, 4000, 1), "general")])
def test_batch_target_in_original_units_equals_host_division(name, maker, path):
    """Trajectories 0-3 take set_target on the host's x̃ ./ D_c, trajectories
    4-7 set_target_original on the same x̃ (the division on the device, the
    pump's src/pump.jl:331): each pair ends alike bit for bit (termination,
    iteration count, ℓ1 gap, stored target, point)."""
    print("batch target original units", name, path, flush=True)
    m = maker()
    pp = projector_problem(m)
    n = m.A.shape[1]
    p = SolverParameters(eps_optimal_relative=1e-4, eps_feasible_relative=1e-4,
                         iteration_limit=100_000)
    bt = _batch(pp, 8)
    try:
        bt.set_kernel_path(path)
        bt.set_column_scale(pp.D_c)
        for rnd in range(2):  # a cold start, then warm-started projections
            targets = [G.binary_targets(m, 40 + 10 * rnd + k) for k in range(4)]
            for k in range(4):
                if rnd == 0:
                    bt.set_x(k, np.zeros(n))
                    bt.set_x(k + 4, np.zeros(n))
                bt.set_target(k, targets[k] / pp.D_c)
                bt.set_target_original(k + 4, targets[k])
            assert bt.solve(p, 1.0) == ["OPTIMAL"] * 8
            for k in range(4):
                a, b = bt.info(k), bt.info(k + 4)
                assert (a.total_count, a.relative_l1_gap, a.l1_distance) == \
                    (b.total_count, b.relative_l1_gap, b.l1_distance), (rnd, k)
                assert np.array_equal(bt.vector(k, "x_tilde"), bt.vector(k + 4, "x_tilde"))
                assert np.array_equal(bt.pdhg_x(k), bt.pdhg_x(k + 4)), (rnd, k)
    finally:
        bt.close()
