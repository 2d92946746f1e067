"""Exact left-to-right sums on the device (hem_exact_sums; the reductions of
Norm, src/hash_align.c:150-164, and of the finalisation, src/EM.c:362-363).

The device computes them in parallel by two methods: chunk programs (runs
whose partial sums stay in one binade, summed exactly on its grid, the
other values one by one) walked in order, and chunks run from windows of
speculated start values with an in-order lookup (HEM_SUM_WINDOWS).  The bar
for both is bitwise equality with the sequential loop, (((+0 + x0) + x1) +
...), for every input, including inputs built to defeat either method
(partial sums that return to zero or hover on a power of two, round-half
ties on every step, huge dynamic range, denormals, signed zeros,
infinities).  The expected values come from numpy's sequential cumulative
sum (np.add.accumulate adds strictly left to right) and from the oracle's
restatement of Norm.
"""
import sys

import numpy as np
import pytest

import hera_amd as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seq_sum(x):
    return np.cumsum(np.concatenate([[0.0], x]))[-1]


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (a.view(np.int64) == b.view(np.int64))))


def cases():
    rng = np.random.default_rng(7)
    out = {}
    for n in (1, 2, 3, 63, 64, 255, 256, 257, 1023, 1024, 1025, 5000,
              200000):
        out["normal_%d" % n] = rng.normal(size=(3, n))
    out["lognormal_wide"] = rng.lognormal(0, 12, size=(2, 100000))
    big = rng.lognormal(0, 3, size=50000)
    canc = np.concatenate([big, -rng.permutation(big)]) + \
        rng.normal(scale=1e-9, size=100000)
    out["return_to_zero"] = np.stack([canc, canc[::-1]])
    out["alternating"] = (np.where(np.arange(60000) % 2 == 0, 1.0, -1.0) *
                          rng.lognormal(0, 1, 60000))[None, :]
    out["denormal"] = (rng.normal(size=(2, 30000)) * 1e-310)
    z = np.zeros((2, 5000))
    z[0, ::3] = -0.0
    z[1, :] = -0.0
    out["signed_zeros"] = z
    steps = rng.normal(size=70000)
    steps[35000:] *= 1e-12
    out["absorbed_tail"] = np.stack([steps, np.abs(steps)])
    inf = rng.normal(size=(3, 4000))
    inf[0, 1500] = np.inf
    inf[1, 100] = np.inf
    inf[1, 3000] = -np.inf
    inf[2, 2500] = np.nan
    out["nonfinite"] = inf
    # every step a round-half tie: from 2^52 (ulp 1) add halves and odd
    # halves, so each result depends on the parity of the running sum
    ties = rng.integers(0, 4, size=(2, 30000)) + 0.5
    ties[:, 0] = 2.0 ** 52
    ties[1, 1::2] *= -1.0
    out["ties"] = ties
    # partial sums hovering on 1.0, a binade boundary, by steps of a few
    # ulps of 1, then far below it
    hov = rng.integers(-3, 4, size=(2, 50000)) * 2.0 ** -53
    hov[:, 0] = 1.0
    hov[1, 25000:] *= 2.0 ** 20
    out["boundary_hover"] = hov
    # partial sums near the top of the double range and near the bottom of
    # the normal range
    out["huge"] = rng.normal(size=(2, 20000)) * 2.0 ** 1010
    out["tiny_normal"] = rng.normal(size=(2, 20000)) * 2.0 ** -1000
    # chunk sums that miss the sequential sum by 1 (1e30 + 1 - 1e30 is 0
    # left to right), so every later chunk's start is guessed 1 too high
    # against running sums of order 1e-3: the guessed runs do not hold for
    # the true starts and the walk adds those values one by one
    mis = rng.normal(size=(2, 50000)) * 1e-3
    mis[:, :3] = [1e30, 1.0, -1e30]
    mis[1, 20000:20003] = [1e30, 1.0, -1e30]
    out["misguided"] = mis
    # short values (a few significant bits, as late SQUAREM r and v rows
    # have: differences of nearby doubles) whose running sums wander
    # around zero across many binades, so runs on a row's grid cross
    # binades; then short values mixed with full-length ones, and values
    # on one coarse grid interleaved with values on finer ones
    walk = rng.integers(-2000, 2001, size=(2, 80000)) * 2.0 ** -60
    walk[1, ::7] *= 2.0 ** -9
    out["short_walk"] = walk
    mixed = rng.integers(-64, 65, size=(2, 60000)) * 2.0 ** -40
    mixed[0, ::5] = rng.normal(size=12000) * 2.0 ** -34
    mixed[1, ::3] *= -(2.0 ** -20)
    out["short_mixed"] = mixed
    return out


ENGINES = {"programs": 0, "windows": H.SUM_WINDOWS}


@pytest.mark.parametrize("engine", sorted(ENGINES))
@pytest.mark.parametrize("name", sorted(cases().keys()))
def test_exact_sums_bitwise(name, engine):
    x = cases()[name]
    e = ENGINES[engine]
    got, reruns = H.exact_sums(x, H.SUM_PLAIN | e)
    exp = np.array([seq_sum(r) for r in x])
    sq, _ = H.exact_sums(x, H.SUM_SQUARES | e)
    exp_sq = np.array([seq_sum(r * r) for r in x])
    say("%-16s %-8s n=%6d rows=%d reruns=%d" % (
        name, engine, x.shape[1], x.shape[0], reruns))
    assert same_bits(got, exp), (got, exp)
    assert same_bits(sq, exp_sq), (sq, exp_sq)


def launch_chunks(rows, n):
    """Chunks of one launch of the signed sums: chains of 256 values up to
    16,384 values per row, of 1,024 above."""
    L = 256 if n <= 16384 else 1024
    return rows * ((n + L - 1) // L)


ADVERSARIAL = ["return_to_zero", "alternating", "denormal", "absorbed_tail",
               "nonfinite", "signed_zeros", "lognormal_wide"]


@pytest.mark.parametrize("name", ADVERSARIAL)
def test_exact_sums_bitwise_narrow_windows(name):
    # Signed sums take 128-candidate windows when a launch holds at most
    # 512 chunks and 64-candidate windows above that: the adversarial rows
    # stacked (each copy scaled by a power of two) until the launch passes
    # 512 chunks run the 64-wide path.
    x = cases()[name]
    reps = 1
    while launch_chunks(reps * x.shape[0], x.shape[1]) <= 512:
        reps += 1
    scale = 2.0 ** (np.arange(reps) % 5 - 2)
    xs = np.concatenate([x * s for s in scale])
    assert launch_chunks(xs.shape[0], xs.shape[1]) > 512
    got, reruns = H.exact_sums(xs, H.SUM_PLAIN | H.SUM_WINDOWS)
    exp = np.array([seq_sum(r) for r in xs])
    sq, _ = H.exact_sums(xs, H.SUM_SQUARES | H.SUM_WINDOWS)
    exp_sq = np.array([seq_sum(r * r) for r in xs])
    say("%-16s n=%6d rows=%d reruns=%d (64-wide)" % (
        name, xs.shape[1], xs.shape[0], reruns))
    assert same_bits(got, exp)
    assert same_bits(sq, exp_sq)


@pytest.mark.parametrize("engine", sorted(ENGINES))
@pytest.mark.parametrize("name", ["normal_1025", "normal_200000",
                                  "lognormal_wide", "return_to_zero",
                                  "alternating", "denormal", "ties",
                                  "boundary_hover"])
def test_norm_sd_bitwise_oracle(name, engine):
    x = cases()[name]
    got, _ = H.exact_sums(x, H.SUM_NORM_SD | ENGINES[engine])
    exp = np.array([O.norm_sd(r) for r in x])
    assert same_bits(got, exp), (got, exp)


def test_norm_sd_on_squarem_vectors(cfg1):
    """r = u1 - u, v = (u2 - u1) - r along the reference trajectory: the
    vectors whose norms set every step length."""
    cls, u = cfg1
    rows = []
    for k in (0, 3, 20, 60):
        lam = (np.full(cls.n_tx, 1.0 / cls.n_tx) if k == 0 else
               O.estimate_count(cls, u, min_round=k, max_round=k)["lam"])
        u1, _ = O.em_map(cls, u, lam)
        u2, _ = O.em_map(cls, u, u1)
        r = u1 - lam
        rows += [r, (u2 - u1) - r]
    x = np.stack(rows)
    exp = np.array([O.norm_sd(r) for r in x])
    for engine, e in sorted(ENGINES.items()):
        got, reruns = H.exact_sums(x, H.SUM_NORM_SD | e)
        say("squarem vectors, %s: reruns=%d over %d rows" % (
            engine, reruns, len(rows)))
        assert same_bits(got, exp)


def test_solve_exact_equals_single_lane(cfg1):
    """Both methods of the parallel exact sums and the single-lane chains
    give the same solve, bit for bit, under both step forms."""
    cls, u = cfg1
    p = H.Problem(cls)
    try:
        for step in (H.STEP_SD, H.STEP_L2):
            b = p.solve(u, norm=H.NORM_SEQ, step=step, max_iter=120)
            for norm in (H.NORM_EXACT, H.NORM_EXACT_WIN):
                a = p.solve(u, norm=norm, step=step, max_iter=120)
                say("step=%d norm=%d k=%d reruns=%d" % (
                    step, norm, a["stats"]["iters"],
                    a["stats"]["sum_reruns"]))
                assert a["stats"]["iters"] == b["stats"]["iters"]
                np.testing.assert_array_equal(a["lam"], b["lam"])
                np.testing.assert_array_equal(a["tpm"], b["tpm"])
    finally:
        p.close()


def test_exact_sums_rejects_bad_arguments():
    with pytest.raises(H.HemError):
        H.exact_sums(np.zeros((1, 4)), 7)
    with pytest.raises(H.HemError):
        H.exact_sums(np.zeros((1, 0)))
