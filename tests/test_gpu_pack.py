"""GPU tests of the block-diagonal pack of independent instances (pdfp_pack_*).

Each instance must behave as its own FP32 projection: its KKT residuals,
re-evaluated in FP64 by the oracle, meet the tolerance (×10 for FP32
rounding), its ℓ1 distance matches the single FP32 projector's to 1e-4
relative, and instances that are not active are left unchanged."""
import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import generators as G
from pdfp_amd import pdfp as P
from pdfp_amd.milp import read_mps
from pdfp_amd.pack import PDFPPack
from pdfp_amd.pdfp import SolverParameters, projector_problem

from .conftest import example_path

pytestmark = pytest.mark.gpu


def _oracle_kkt(pp, xt_scaled, x, y, z, ds):
    o = O.OraclePDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w)
    O.set_target(o, xt_scaled)
    o.initialize(None, 1.0)
    return o.residual_of(x, y, z, ds)


def _single(pp, target, p, pw, D_r=None):
    s = P.PDFPState(pp.A, pp.b, pp.is_eq, pp.is_int, pp.var_lb, pp.var_ub, pp.w,
                    precision="fp32")
    if D_r is not None:
        P.original_primal_residual(s, D_r)
    P.set_target(s, target)
    assert P.pdfp_solve(s, p, pw) == "OPTIMAL"
    l1 = s.info().l1_distance
    s.close()
    return l1


def _instances():
    # uneven sizes, both pure-binary and mixed (continuous columns, w = 0)
    return [G.set_cover(133, 201, 1), G.mixed(400, 800, 3), G.set_cover(700, 1500, 4),
            G.mixed(60, 90, 5), G.set_cover(2000, 3500, 6)]


def test_pack_instances_are_independent_projections():
    print("pack independent", flush=True)
    milps = _instances()
    pps = [projector_problem(m) for m in milps]
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    pk = PDFPPack(pps)
    targets = [G.binary_targets(m, 50 + k) / pp.D_c for k, (m, pp) in enumerate(zip(milps, pps))]
    for k, t in enumerate(targets):
        pk.set_target(k, t)
    pw = np.linspace(0.7, 1.4, len(pps))
    assert pk.solve(p, pw) == ["OPTIMAL"] * len(pps)
    for k, pp in enumerate(pps):
        info = pk.info(k)
        assert info.total_count > 0 and info.step_size > 0
        r = _oracle_kkt(pp, targets[k], pk.pdhg_x(k), pk.vector(k, "pdhg_y"),
                        pk.vector(k, "pdhg_z"), pk.vector(k, "dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
        l1s = _single(pp, targets[k], p, float(pw[k]))
        assert abs(info.l1_distance - l1s) <= 1e-4 * max(1.0, l1s), (k, info.l1_distance, l1s)


def test_pack_inactive_instance_is_untouched_and_warm_starts():
    print("pack inactive", flush=True)
    milps = [G.set_cover(133, 201, 4), G.set_cover(150, 260, 7), G.mixed(80, 120, 2)]
    pps = [projector_problem(m) for m in milps]
    p = SolverParameters(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
    pk = PDFPPack(pps)
    for k, pp in enumerate(pps):
        pk.set_target(k, np.zeros(pp.A.shape[1]))
    assert pk.solve(p, 1.0) == ["OPTIMAL"] * 3
    before = [pk.pdhg_x(k) for k in range(3)]
    n1 = pk.total_count(1)
    pk.set_target(0, np.ones(pps[0].A.shape[1]) / pps[0].D_c)  # feasible target: distance 0
    terms = pk.solve(p, 1.0, active=[1, 0, 0])
    assert terms[0] == "OPTIMAL"
    assert pk.info(0).l1_distance <= 1e-5
    for k in (1, 2):
        assert np.array_equal(pk.pdhg_x(k), before[k])
    assert pk.total_count(1) == n1


def test_pack_original_primal_residual():
    """original_primal_residual! on one mixed instance of the pack: the
    primal residual in original units, as the single projector computes it."""
    print("pack orig primal", flush=True)
    milps = [G.mixed(300, 600, 8), G.set_cover(200, 300, 9)]
    pps = [projector_problem(m) for m in milps]
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    pk = PDFPPack(pps)
    pk.original_primal_residual(0, pps[0].D_r)
    targets = [G.binary_targets(m, 11 + k) / pp.D_c for k, (m, pp) in enumerate(zip(milps, pps))]
    for k, t in enumerate(targets):
        pk.set_target(k, t)
    assert pk.solve(p, 1.0) == ["OPTIMAL", "OPTIMAL"]
    for k, pp in enumerate(pps):
        l1s = _single(pp, targets[k], p, 1.0, D_r=pp.D_r if k == 0 else None)
        l1p = pk.info(k).l1_distance
        assert abs(l1p - l1s) <= 1e-4 * max(1.0, l1s), (k, l1p, l1s)


def test_pack_block_graph_replays_equal_stream_launches():
    """Blocks replayed as captured graphs give the same projections, bit for
    bit, as the same launches on the stream, with instances that end and
    start again at different blocks.  Below the autotuner's threshold (15k
    nonzeros in all) every pack gets the same schedule."""
    print("pack block graph", flush=True)
    milps = [G.set_cover(133, 201, 1), G.mixed(300, 600, 3), G.mixed(60, 90, 5),
             G.set_cover(400, 700, 4)]
    pps = [projector_problem(m) for m in milps]
    loose = SolverParameters(eps_optimal_relative=1e-3, eps_feasible_relative=1e-4)
    tight = SolverParameters(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6,
                             iteration_limit=200_000)
    runs = []
    for graph in (False, True, True):
        pk = PDFPPack(pps)
        pk.set_kernel_path("general")   # every instance fits in LDS
        pk.set_block_graph(graph)
        rec = []
        restarted = [False] * len(pps)
        for k, (m, pp) in enumerate(zip(milps, pps)):
            pk.set_target(k, G.binary_targets(m, 20 + k) / pp.D_c)
            pk.begin(k, tight if k % 2 else loose, 0.8 + 0.2 * k)
        running = [True] * len(pps)
        while any(running):
            terms, _ = pk.run(until_any=True)
            for k in range(len(pps)):
                if not running[k] or terms[k] == "UNSPECIFIED":
                    continue
                assert terms[k] == "OPTIMAL"
                info = pk.info(k)
                assert info.block_graph == int(graph)
                rec.append((k, info.total_count, info.n_restarts, info.l1_distance,
                            pk.pdhg_x(k), pk.vector(k, "pdhg_z")))
                running[k] = False
                if not restarted[k]:       # each instance once more, to a new target
                    restarted[k] = True
                    pk.set_target(k, G.binary_targets(milps[k], 40 + k) / pps[k].D_c)
                    pk.begin(k, loose, 1.0)
                    running[k] = True
        pk.close()
        runs.append(rec)
    assert len(runs[0]) == 2 * len(pps)
    for run in runs[1:]:
        assert len(run) == len(runs[0])
        for a, b in zip(runs[0], run):
            assert a[:4] == b[:4]
            assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


def test_pack_small_instances_equal_single_projector_bitwise():
    """When every instance fits in LDS, the pack runs one workgroup per
    running instance, each iterating as the single FP32 projector's
    one-workgroup path: every instance's projections (a cold one, then a
    warm-started one begun at another block) equal the single projector's
    bit for bit, and the general passes (path forced) reach the same ℓ1
    distances to the tolerance."""
    print("pack small", flush=True)
    milps = [read_mps(example_path("p0201")), read_mps(example_path("pk1")),
             G.set_cover(133, 201, 2), G.mixed(60, 90, 5), read_mps(example_path("markshare2"))]
    pps = [projector_problem(m) for m in milps]
    tol = 1e-6
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=100_000)
    pw = [1.0 + 0.1 * k for k in range(len(pps))]
    t1 = [G.binary_targets(m, 60 + k) / pp.D_c for k, (m, pp) in enumerate(zip(milps, pps))]
    t2 = [G.binary_targets(m, 80 + k) / pp.D_c for k, (m, pp) in enumerate(zip(milps, pps))]
    # the single projector, one instance at a time
    ref = []
    for k, pp in enumerate(pps):
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
    pk = PDFPPack(pps)
    for k in range(len(pps)):
        pk.set_target(k, t1[k])
        pk.begin(k, p, pw[k])
    running, second = [True] * len(pps), [False] * len(pps)
    got = [[] for _ in pps]
    while any(running):
        # one block per call: the pump's form while some instances round
        terms, _ = pk.run(until_any=True, one_block=True)
        for k in range(len(pps)):
            if not running[k] or terms[k] == "UNSPECIFIED":
                continue
            assert terms[k] == "OPTIMAL" and pk.info(k).small_path == 1
            got[k].append((pk.total_count(k), pk.info(k).l1_distance, pk.pdhg_x(k)))
            running[k] = False
            if not second[k]:
                second[k] = running[k] = True
                pk.set_target(k, t2[k])
                pk.begin(k, p, pw[k])
    for k in range(len(pps)):
        for a, b in zip(ref[k], got[k]):
            assert a[0] == b[0] and a[1] == b[1], (k, a[:2], b[:2])
            assert np.array_equal(a[2], b[2]), k
    assert pk._lib.pdfp_pack_run(pk._ctx, 3, None, None, None) != 0   # no such mode
    pk.set_kernel_path("general")
    for k in range(len(pps)):
        pk.set_target(k, t1[k])
    assert pk.solve(p, np.asarray(pw)) == ["OPTIMAL"] * len(pps)
    for k in range(len(pps)):
        assert pk.info(k).small_path == 0
        l1g, l1s = pk.info(k).l1_distance, ref[k][0][1]
        assert abs(l1g - l1s) <= 1e-4 * max(1.0, l1s), (k, l1g, l1s)


@pytest.mark.parametrize("path", ["auto", "general"])
def test_pack_instance_with_empty_row_and_column(path):
    """An instance with an empty row (0 ≤ 1), an equality row and an empty
    column with an infinite upper bound, packed beside a set cover, on both
    kernel paths: each instance ends as the single FP32 projector does on
    it alone (termination, ℓ1 distance), and the oracle's FP64 residuals of
    the pack's point meet the tolerance."""
    import scipy.sparse as sp

    print("pack edges", path, flush=True)
    A = sp.csc_matrix(np.array([[1.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0],
                                [0.0, 1.0, 1.0, 0.0]]))
    edge = G.MILP(c=np.zeros(4), c0=0.0, A=A, b=np.array([1.0, 1.0, 1.0]),
                  row_sense=np.array(["le", "le", "eq"]), l=np.zeros(4),
                  u=np.array([1.0, 1.0, 1.0, np.inf]), int_idx=np.array([0, 1, 2]))
    milps = [edge, G.set_cover(133, 201, 3)]
    pps = [projector_problem(m) for m in milps]
    tol = 1e-5
    p = SolverParameters(eps_optimal_relative=tol, eps_feasible_relative=tol,
                         iteration_limit=200_000)
    targets = [np.array([1.0, 1.0, 1.0, 5.0]) / pps[0].D_c,
               G.binary_targets(milps[1], 7) / pps[1].D_c]
    pk = PDFPPack(pps)
    pk.set_kernel_path(path)
    for k, t in enumerate(targets):
        pk.set_target(k, t)
    assert pk.solve(p, np.ones(2)) == ["OPTIMAL"] * 2
    for k, pp in enumerate(pps):
        assert pk.info(k).small_path == (1 if path == "auto" else 0)
        r = _oracle_kkt(pp, targets[k], pk.pdhg_x(k), pk.vector(k, "pdhg_y"),
                        pk.vector(k, "pdhg_z"), pk.vector(k, "dual_slack"))
        assert r[0] <= 10 * tol and r[1] <= 10 * tol and r[2] <= 10 * tol, (k, r)
        l1s = _single(pp, targets[k], p, 1.0)
        l1 = pk.info(k).l1_distance
        assert abs(l1 - l1s) <= 1e-4 * max(1.0, l1s), (k, l1, l1s)
    pk.close()
