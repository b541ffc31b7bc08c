"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/pdfp_hip.h declares, and its host-only logic (defaults, argument
validation, the power-iteration start vector) behaves.  No device compute."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from pdfp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pdfp_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b((?:pdfp|pdlp)_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name


def test_default_params_are_the_reference_defaults():
    """src/pdlp.jl:10-31."""
    p = _lib.PdfpParams()
    _lib.load().pdfp_default_params(C.byref(p))
    assert (p.eps_optimal_relative, p.eps_feasible_relative) == (1e-4, 1e-4)
    assert p.iteration_limit == 100_000 and p.time_limit == 3600.0
    assert p.termination_evaluation_frequency == 200 and p.reflection_coefficient == 1.0
    assert p.sv_max_iterations == 5000 and p.sv_tolerance == 1e-4
    assert (p.artificial_restart_threshold, p.sufficient_reduction_for_restart,
            p.necessary_reduction_for_restart) == (0.36, 0.2, 0.5)
    assert (p.k_p, p.k_i, p.k_d, p.i_smooth) == (0.99, 0.01, 0.0, 0.3)


def test_start_vector_matches_oracle_bitwise():
    lib = _lib.load()
    for m in (1, 7, 1000):
        out = np.empty(m)
        lib.pdfp_start_vector(m, 20240517, _lib.dptr(out))
        assert np.array_equal(out, O.start_vector(m))


def test_create_rejects_bad_arguments_before_touching_a_device():
    lib = _lib.load()
    ctx = C.c_void_p()
    colptr = np.array([0, 1, 2], dtype=np.int64)
    rowval = np.array([0, 0], dtype=np.int64)
    vals = np.array([1.0, 1.0])
    b = np.array([1.0])
    eq = np.zeros(1, dtype=np.uint8)
    ii = np.ones(2, dtype=np.uint8)
    lb, ub, w = np.zeros(2), np.ones(2), np.ones(2)
    args = [C.byref(ctx), 1, 2, 2, _lib.i64ptr(colptr), _lib.i64ptr(rowval), _lib.dptr(vals), 0,
            _lib.dptr(b), _lib.u8ptr(eq), _lib.u8ptr(ii), _lib.dptr(lb), _lib.dptr(ub),
            _lib.dptr(w), 0, 0]
    bad = list(args)
    bad[3] = 3                                  # nnz disagrees with colptr
    assert lib.pdfp_create(*bad) == 1
    assert b"nnz" in lib.pdfp_last_error(None)
    bad = list(args)
    bad[14] = 7                                 # unknown precision
    assert lib.pdfp_create(*bad) == 1
    bad = list(args)
    bad[7] = 2                                  # index base
    assert lib.pdfp_create(*bad) == 1
    bad = list(args)
    bad[8] = None                               # null b
    assert lib.pdfp_create(*bad) == 1
    assert not ctx.value
    # calls on a null context fail cleanly
    assert lib.pdfp_warm_start(None) == 1
    lib.pdfp_destroy(None)


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_pack_create_rejects_bad_arguments_before_touching_a_device():
    lib = _lib.load()
    ctx = C.c_void_p()
    colptr = np.array([0, 1, 2], dtype=np.int64)
    rowval = np.array([0, 0], dtype=np.int64)
    vals = np.array([1.0, 1.0])
    b = np.array([1.0])
    eq = np.zeros(1, dtype=np.uint8)
    ii = np.ones(2, dtype=np.uint8)
    lb, ub, w = np.zeros(2), np.ones(2), np.ones(2)
    m = np.array([1, 1], dtype=np.int64)
    n = np.array([2, 2], dtype=np.int64)

    def arr(ct, a):
        out = (C.POINTER(ct) * 2)()
        out[0] = out[1] = a.ctypes.data_as(C.POINTER(ct))
        return out

    args = [C.byref(ctx), 2, m.ctypes.data_as(C.POINTER(C.c_int64)),
            n.ctypes.data_as(C.POINTER(C.c_int64)), arr(C.c_int64, colptr),
            arr(C.c_int64, rowval), arr(C.c_double, vals), 0, arr(C.c_double, b),
            arr(C.c_uint8, eq), arr(C.c_uint8, ii), arr(C.c_double, lb), arr(C.c_double, ub),
            arr(C.c_double, w), 0]
    bad = list(args)
    bad[1] = 0                                  # no instance
    assert lib.pdfp_pack_create(*bad) == 1
    bad = list(args)
    bad[7] = 2                                  # index base
    assert lib.pdfp_pack_create(*bad) == 1
    bad = list(args)
    bad[8] = None                               # null b array
    assert lib.pdfp_pack_create(*bad) == 1
    neg = np.array([1, -1], dtype=np.int64)
    bad = list(args)
    bad[2] = neg.ctypes.data_as(C.POINTER(C.c_int64))
    assert lib.pdfp_pack_create(*bad) == 1
    assert b"negative" in lib.pdfp_pack_last_error(None)
    assert not ctx.value
    assert lib.pdfp_pack_warm_start(None, 0) == 1
    lib.pdfp_pack_destroy(None)


def test_uneven_set_covers_follow_their_recipe():
    """Sizes e^U[ln lo, ln hi] per seed, each instance a set cover of them."""
    from pdfp_amd import generators as G

    ms = G.uneven_set_covers(3, m_range=(20, 60), n_range=(40, 120), first_seed=5)
    for s, m in zip(range(5, 8), ms):
        rng = np.random.Generator(np.random.PCG64(s))
        mm = int(round(float(np.exp(rng.uniform(np.log(20), np.log(60))))))
        nn = int(round(float(np.exp(rng.uniform(np.log(40), np.log(120))))))
        assert m.A.shape == (mm, nn)
        assert np.array_equal(m.A.toarray(), G.set_cover(mm, nn, s).A.toarray())


@pytest.mark.parametrize("workload,keys", [
    ("mixed-1m", ("EpiBatchPrimalY<8, false>", "EpiBatchDualZ<8, false>",
                  "EpiPrimalY<float, false>", "EpiDualZ<float, false>")),
    ("powerlaw-5m", ("EpiBatchPrimalYNT<8, false>", "EpiSlicePartialLanes<8>",
                     "EpiBatchDualZLongNT<8, false>",
                     "EpiPrimalY<float, false>", "EpiDualZ<float, false>")),
    ("general-int-4m", ("EpiBatchPrimalY<8, false>", "EpiSlicePartial<8>",
                        "k_batch_slice_dualz<8, false>",
                        "EpiPrimalY<float, false>", "EpiDualZ<float, false>")),
    ("setcover-pack-64", ("EpiPackPrimalY<false>", "EpiPackDualZ<false>")),
])
def test_bench_finds_the_traffic_of_each_workload(workload, keys):
    """bench.py's roofline.traffic comes from the committed counter summary
    of its workload: every workload name reaches its file and the passes it
    may report as dominant."""
    import bench

    for key in keys:
        traffic = bench.measured_traffic(workload, key)
        assert traffic is not None and traffic > 0, (workload, key)


def _julia_struct_fields(name: str):
    """(field, type) pairs of a `struct name ... end` block in INTEGRATION.md's shim."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = text.split(f"struct {name}", 1)[1].split("\nend", 1)[0]
    body = body.split("\n", 1)[1]                      # drop the comment on the header line
    return re.findall(r"(\w+)::(Float64|Int64|Int32)", body)


def _c_type_name(ctype):
    return {C.c_double: "Float64", C.c_int64: "Int64", C.c_int32: "Int32"}[ctype]


@pytest.mark.parametrize("julia,ctype", [("SolverParametersC", _lib.PdfpParams),
                                         ("PDFPInfoC", _lib.PdfpInfo),
                                         ("PDLPInfoC", _lib.PdlpInfo)])
def test_julia_shim_structs_mirror_the_c_structs(julia, ctype):
    """The Julia structs of the integration shim have the C structs' fields in
    the same order with the same types, so `Ref{…}` passes them by layout."""
    want = [(f, _c_type_name(t)) for f, t in ctype._fields_]
    assert _julia_struct_fields(julia) == want


def test_probe_library_is_separate():
    """The benchmark's gather probe lives in libpdfp_probe.so (its own header,
    include/pdfp_probe.h), not in the product library."""
    src = open(os.path.join(ROOT, "include", "pdfp_probe.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(pdfp_[a-z0-9_]+)\s*\(", src)))
    assert names == ["pdfp_gather_ceiling", "pdfp_probe_last_error"]
    probe = _lib.load_probe()
    lib = _lib.load()
    for name in names:
        assert hasattr(probe, name)
        assert not hasattr(lib, name), name
