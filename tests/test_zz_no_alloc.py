"""GPU test of the allocation-free entry point pdfp_max_violation_noalloc.

It must return what the allocating entry point pdfp_checker_max_violation
returns for the same model and point (the same kernels: bit-identical), on
two calls in a row, the second on a smaller model, so that the shared
static scratch array is reused with fewer partials; and it must refuse, from
the host and before any launch, a row count beyond its scratch array.

Device buffers for the inputs and the output are made here with the HIP
runtime that the library itself links (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from pdfp_amd import _lib
from pdfp_amd import generators as G
from pdfp_amd.milp import DeviceViolationCheck

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        _lib.load()                     # the library's HIP runtime is loaded first
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipDeviceSynchronize()
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(hip, milp, x):
    A = sp.csr_matrix(milp.A, dtype=np.float64)
    A.sort_indices()
    m, n = A.shape
    args = [hip.put(np.asarray(A.indptr, dtype=np.int32)),
            hip.put(np.asarray(A.indices, dtype=np.int32)),
            hip.put(np.asarray(A.data, dtype=np.float64)),
            hip.put(np.asarray(milp.b, dtype=np.float64)),
            hip.put(np.asarray(milp.row_sense == "le", dtype=np.uint8)),
            hip.put(np.asarray(x, dtype=np.float64))]
    out = hip.put(np.full(1, np.nan))
    lib = _lib.load()
    st = lib.pdfp_max_violation_noalloc(m, n, *args, out, None)
    assert st == 0, lib.pdfp_checker_last_error(None)
    return float(hip.get(out, np.zeros(1))[0])


def test_noalloc_max_violation_equals_the_checker_on_two_calls():
    hip = _Hip()
    try:
        big = G.mixed(3000, 6000, 2)                       # 12 workgroups of partials
        small = G.general_integer(300, 700, 3)             # 2 workgroups, equality rows
        for k, milp in enumerate((big, small)):
            print("noalloc call", k, milp.A.shape, flush=True)
            rng = np.random.default_rng(40 + k)
            x = np.round(rng.uniform(milp.l, np.minimum(milp.u, milp.l + 3.0)))
            want = DeviceViolationCheck(milp)(x)
            got = _noalloc(hip, milp, x)
            assert got == want, (k, got, want)
            assert want > 0.0
    finally:
        hip.free()


def test_noalloc_rejects_more_rows_than_its_scratch():
    lib = _lib.load()
    dummy = C.c_void_p(8)    # never dereferenced: the size check comes first
    st = lib.pdfp_max_violation_noalloc(16_777_217, 10, dummy, dummy, dummy, dummy, dummy,
                                        dummy, dummy, None)
    assert st == 1
    assert b"exceeds the maximum" in lib.pdfp_checker_last_error(None)
