"""The oracle's restatement of hera's bootstrap sampler (create_distribution
and bootstrap_sample, src/EM.c:98-150), pinned against the system C
library's rand_r and a direct transcription of the reference loop that
calls it.  Host only.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

LIBC = ctypes.CDLL("libc.so.6")
LIBC.rand_r.restype = ctypes.c_int
LIBC.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]


@pytest.mark.parametrize("seed", [0, 1, 7, 12345, 2 ** 31 - 1, 2 ** 31,
                                  2 ** 32 - 1])
def test_rand_r_equals_libc(seed):
    vals, fin = O.rand_r_seq(seed, 20000)
    s = ctypes.c_uint(seed)
    ref = np.array([LIBC.rand_r(ctypes.byref(s)) for _ in range(20000)],
                   dtype=np.uint32)
    np.testing.assert_array_equal(vals, ref)
    assert fin == s.value
    assert vals.max() < 2 ** 31


def reference_loop(unique, seed):
    """src/EM.c:98-113 and :133-135 with the C library's rand_r."""
    table = np.repeat(np.arange(len(unique)), unique)
    count = np.zeros(len(unique), np.uint32)
    s = ctypes.c_uint(seed)
    n = len(table)
    while True:
        n -= 1
        if n <= 0:
            break
        count[table[LIBC.rand_r(ctypes.byref(s)) % n]] += 1
    return count, s.value


@pytest.mark.parametrize("seed", [0, 99])
def test_sampler_equals_reference_loop(seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 40, 120).astype(np.uint32)
    u[:5] = 0
    got, fin = O.hera_sample(u, 3, seed)
    s = seed
    for r in range(3):
        ref, s = reference_loop(u, s)
        np.testing.assert_array_equal(got[r], ref)
    assert fin == s


def test_sampler_properties():
    rng = np.random.default_rng(5)
    u = rng.integers(0, 1000, 2000).astype(np.uint32)
    u[rng.random(2000) < 0.3] = 0
    N = int(u.sum())
    got, _ = O.hera_sample(u, 4, 0)
    assert (got.sum(axis=1) == N - 1).all()      # N - 1 draws
    assert (got[:, u == 0] == 0).all()           # no reads, no draws
    # the shrinking range favours early slots: the first transcripts'
    # mean share exceeds their read share (a property of the reference's
    # sampler, not of a multinomial)
    cum = np.cumsum(u)
    head = cum < N // 2
    assert got[:, head].sum() / got.sum() > u[head].sum() / N
    one, _ = O.hera_sample(np.array([0, 1, 0], np.uint32), 2, 3)
    assert (one == 0).all()                      # N = 1: no draws
    none, fin = O.hera_sample(np.zeros(4, np.uint32), 2, 3)
    assert (none == 0).all() and fin == 3
