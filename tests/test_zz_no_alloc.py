"""hem_exact_sums_noalloc: the exact left-to-right sums with the library's
fixed device scratch (no allocation), called through the C ABI on device
rows; every result bitwise hem_exact_sums's (by the same method, windows of
speculated starts, with the same reruns, and by the default method, chunk
programs), on two calls in a row (the second smaller, so stale scratch from
the first would show).

Device memory and streams for the rows come from the HIP runtime the
library itself is linked against (found among the process's mappings), so
the test runs in any process that has loaded the library."""
import ctypes as C
import sys

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip_runtime():
    """The libamdhip64 mapped into this process (the library's own)."""
    _lib.load()
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64" in p:
                paths.add(p)
    assert len(paths) == 1, paths
    rt = C.CDLL(paths.pop())
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    rt.hipStreamDestroy.argtypes = [C.c_void_p]
    rt.hipStreamSynchronize.argtypes = [C.c_void_p]
    return rt


class Dev:
    """A device buffer of `nbytes` from the library's HIP runtime."""

    def __init__(self, rt, nbytes):
        self.rt = rt
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.rt.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes,
                                 D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def noalloc(rt, x, mode, stream):
    lib = _lib.load()
    rows, n = x.shape
    xd, od, rd = Dev(rt, x.nbytes), Dev(rt, rows * 8), Dev(rt, 4)
    try:
        xd.put(x.astype(np.float64))
        od.put(np.full(rows, -1.0))
        rd.put(np.array([12345], np.uint32))
        _lib.check(lib.hem_exact_sums_noalloc(xd.p, n, rows, mode, od.p,
                                              rd.p, stream))
        assert rt.hipStreamSynchronize(stream) == 0
        return od.get(np.float64, rows), int(rd.get(np.uint32, 1)[0])
    finally:
        for d in (xd, od, rd):
            d.free()


def rows_big():
    rng = np.random.default_rng(21)
    big = rng.lognormal(0, 3, size=100000)
    canc = np.concatenate([big, -rng.permutation(big)]) + \
        rng.normal(scale=1e-9, size=200000)
    return np.stack([rng.normal(size=200000), canc, canc[::-1],
                     rng.lognormal(0, 8, 200000)])


def rows_small():
    rng = np.random.default_rng(22)
    x = rng.normal(size=(3, 5000))
    x[1, 10] = -0.0
    x[2] *= 1e-310
    return x


def test_noalloc_sums_equal_the_allocating_entry_point():
    rt = hip_runtime()
    side = C.c_void_p()
    assert rt.hipStreamCreate(C.byref(side)) == 0
    try:
        for mode in (H.SUM_PLAIN, H.SUM_NORM_SD, H.SUM_SQUARES):
            # a large call on the default stream, then a smaller one on
            # a stream of its own
            for x, stream in ((rows_big(), None), (rows_small(), side)):
                got, rr = noalloc(rt, x, mode, stream)
                # the same method (windows of speculated starts) on
                # allocated scratch: same bits, same chunks run again
                exp, rr_exp = H.exact_sums(x, mode | H.SUM_WINDOWS)
                say("mode %d rows %d n %d reruns %d/%d" % (
                    mode, x.shape[0], x.shape[1], rr, rr_exp))
                np.testing.assert_array_equal(got.view(np.int64),
                                              exp.view(np.int64))
                assert rr == rr_exp
                # and the default method (chunk programs): same bits
                prog, _ = H.exact_sums(x, mode)
                np.testing.assert_array_equal(got.view(np.int64),
                                              prog.view(np.int64))
    finally:
        rt.hipStreamDestroy(side)


def test_noalloc_sums_reject_sizes_beyond_the_scratch():
    # 5 rows of 10^6 values are 4,885 chunks of 1,024: over 4,096
    with pytest.raises(H.HemError) as ei:
        _lib.check(_lib.load().hem_exact_sums_noalloc(
            C.c_void_p(8), 1_000_000, 5, H.SUM_PLAIN, C.c_void_p(8), None,
            None))
    assert ei.value.code == _lib.HEM_ERR_RANGE
