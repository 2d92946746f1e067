"""Pins the bootstrap resampler restatement (oracle) and the tiling rules.

The resampler draws Multinomial(N, w/N) over all fragments (unique reads and
class reads), which is hera's -b bootstrap (src/EM.c:123-150) with the class
counts resampled as well (hera itself only resamples unique counts and its
shared rand_r seed races across threads, so its replicates cannot be
reproduced).  Checks: the Philox-4x32-10 known-answer vectors of the
published generator, exact conservation of N, the multinomial law (exact
pmf on a tiny case, first two moments on a large one), determinism per seed,
and the chunk plan's invariants.
"""
import math

import numpy as np
import pytest

import hera_amd as H
from conftest import tiny_problem
from oracle import oracle as O


@pytest.mark.parametrize("key,ctr,expect", [
    (0, [0, 0, 0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    (0xffffffffffffffff, [0xffffffff] * 4,
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    (0x299f31d0a4093822, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(key, ctr, expect):
    # Random123 known-answer vectors for philox4x32-10 (key words
    # little-endian in the 64-bit key)
    assert list(O.philox(key, ctr)) == expect


def test_exact_pmf_tiny():
    # N = 4 fragments over weights (1, 1, 2): every outcome's frequency
    # against the multinomial pmf (chi-square, 3000 seeds)
    w = np.array([1, 1, 2], dtype=np.uint32)
    N = int(w.sum())
    p = w / N
    counts = {}
    S = 3000
    for s in range(S):
        c = tuple(O.resample(w, s))
        assert sum(c) == N
        counts[c] = counts.get(c, 0) + 1
    chi2 = 0.0
    n_cells = 0
    for a in range(N + 1):
        for b in range(N + 1 - a):
            c = (a, b, N - a - b)
            pmf = (math.factorial(N) / (math.factorial(a) * math.factorial(b)
                                        * math.factorial(c[2]))
                   * p[0] ** a * p[1] ** b * p[2] ** c[2])
            e = S * pmf
            chi2 += (counts.get(c, 0) - e) ** 2 / e
            n_cells += 1
    # 14 degrees of freedom: P(chi2 > 36) ~ 1e-3
    assert chi2 < 36, chi2


def test_moments_and_conservation():
    rng = np.random.default_rng(0)
    w = rng.integers(0, 500, 3000).astype(np.uint32)
    w[::97] = rng.integers(5000, 20000, len(w[::97]))   # heavy categories
    N = int(w.sum())
    p = w / N
    S = 40
    X = np.array([O.resample(w, 1000 + s) for s in range(S)], dtype=float)
    assert np.all(X.sum(axis=1) == N)
    assert np.all(X[:, w == 0] == 0)
    # mean: z-scores of the per-category means
    m = X.mean(axis=0)
    sd = np.sqrt(N * p * (1 - p) / S)
    z = (m - N * p)[w > 0] / sd[w > 0]
    assert abs(z.mean()) < 0.1 and 0.8 < z.std() < 1.2
    # variance of the heavy categories ~ N p (1 - p)
    hv = w > 4000
    ratio = X[:, hv].var(axis=0, ddof=1) / (N * p[hv] * (1 - p[hv]))
    assert 0.7 < ratio.mean() < 1.3
    # same seed, same counts; another seed, other counts
    np.testing.assert_array_equal(O.resample(w, 7), O.resample(w, 7))
    assert not np.array_equal(O.resample(w, 7), O.resample(w, 8))


def test_chunk_plan_invariants():
    rng = np.random.default_rng(1)
    w = rng.integers(0, 100, 20000).astype(np.uint32)
    w[5::1000] = 10 ** 6
    start, cw = O.resample_plan(w)
    N = int(w.sum())
    Wt = -(-N // 2048)
    assert start[0] == 0 and start[-1] == len(w) and cw[-1] == N
    sizes = np.diff(start.astype(np.int64))
    assert sizes.min() >= 1 and sizes.max() <= 2048
    wsum = np.add.reduceat(w.astype(np.int64), start[:-1].astype(np.int64))
    np.testing.assert_array_equal(np.diff(cw.astype(np.int64)), wsum)
    heavy = np.nonzero(w > Wt)[0]
    for k in heavy:
        j = np.searchsorted(start, k, side="right") - 1
        assert start[j] == k and start[j + 1] == k + 1
    light = wsum[sizes > 1]
    assert light.max() <= Wt


def test_tiling_of_generator_tables(cfg1):
    cls, _ = cfg1
    t = cls.tiling()
    assert t["n_tiles"] > 0 and t["max_tx"] <= 64 and t["max_cls"] <= 320
    # a chain of classes joining 100 transcripts is one component too large
    # for a tile: the table reports no tiling
    members = [[i, i + 1] for i in range(99)]
    big = H.Classes.from_lists(100, members, [1.0] * 99, [500] * 100)
    assert big.tiling()["n_tiles"] == 0
    small, _ = tiny_problem(3, n_tx=30, n_cls=60)
    assert small.tiling()["n_tiles"] >= 1
