"""The batched workloads at their full sizes.

Bootstrap (hera's -b, src/EM.c:233-261): 200,000 transcripts and 2,000,000
drawn classes (60M fragments, seed 0), 100 replicates with seeds 1..100,
each a multinomial redraw of all fragments, solved to a 1e-8 relative
change or 1,000 iterations in one batched run.

Ragged multi-sample batch: 64 samples with transcript counts log-uniform in
5,000-200,000, 10x as many drawn classes, fragments uniform in 5M-80M, 20
replicates each, in one multi-sample handle, plus every sample's gradient
in one call on the same handle.

Every replicate of the default mode (FP64, exact sums) is the reference EM
of its own counts, bit for bit; the oracle checks a few replicates chosen
to include the fastest and slowest to converge (an oracle solve of a 200k
table takes up to ~40 s on one core), and the rest are held to
size-independent properties: fragments conserved by every redraw, counts a
function of the seed alone, identical reruns, each sample's instances
equal to its own handle's.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import hera_amd as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STOP = dict(stop_rel_change=1e-8, min_iter=50, max_iter=1000)
ORA_STOP = dict(stop_rel=1e-8, min_round=50, max_round=1000)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def human():
    return H.generate(200000, 2000000, 60_000_000, seed=0)


def test_bootstrap_100_replicates_200k(human):
    cls, u = human
    R = 100
    seeds = np.arange(R, dtype=np.uint64) + 1
    N = int(u.sum()) + int(cls.cls_count.sum())
    p = H.Problem(cls)
    try:
        say("bootstrap: T=%d C=%d Z=%d N=%d R=%d" % (cls.n_tx, cls.n_cls,
                                                    cls.nnz, N, R))
        b = p.bootstrap(u, seeds, **STOP)
        it = b["iters"]
        say("iterations %d..%d, %d converged, %.1f ms solve" % (
            it.min(), it.max(), b["stats"]["n_converged"],
            b["stats"]["ms_solve"]))
        assert b["est"].shape == (R, cls.n_tx)
        assert np.all((it >= 50) & (it <= 1000))
        # a second run is the same, replicate by replicate
        b2 = p.bootstrap(u, seeds, **STOP)
        np.testing.assert_array_equal(b2["iters"], it)
        np.testing.assert_array_equal(b2["est"], b["est"])
        del b2
        # every redraw keeps all N fragments; counts depend on the seed
        # alone (four launches of 25 seeds against the 100-seed run)
        uo = np.empty((R, cls.n_tx), np.uint32)
        co = np.empty((R, cls.n_cls))
        for a in range(0, R, 25):
            uo[a:a + 25], co[a:a + 25] = p.bootstrap_counts(u, seeds[a:a + 25])
        tot = uo.sum(1, dtype=np.int64) + co.sum(1).astype(np.int64)
        assert np.all(tot == N)
        # est is lambda * len, so sum_t est_t / len_t * len_t over every
        # replicate is finite and positive
        assert np.all(np.isfinite(b["est"])) and np.all(b["est"] >= 0)
        np.testing.assert_allclose(b["tpm"].sum(1), 1e6, rtol=1e-9)
    finally:
        p.close()
    # the oracle on four replicates: fastest, slowest, and two between
    order = np.argsort(it, kind="stable")
    pick = sorted({int(order[0]), int(order[-1]), int(order[R // 2]), 0})
    w = np.concatenate([u.astype(np.uint32),
                        cls.cls_count.astype(np.uint32)])
    for r in pick:
        cnt = O.resample(w, int(seeds[r]))
        np.testing.assert_array_equal(cnt[:cls.n_tx], uo[r])
        np.testing.assert_array_equal(cnt[cls.n_tx:], co[r])
    est, ks = O.bootstrap_batch(cls, uo[pick], co[pick], len(pick),
                                **ORA_STOP)
    say("oracle on replicates %s: k = %s (GPU %s)" % (
        pick, ks.tolist(), it[pick].tolist()))
    np.testing.assert_array_equal(ks, it[pick])
    for j, r in enumerate(pick):
        np.testing.assert_array_equal(b["est"][r], est[j])


def ragged_samples(S=64):
    rng = np.random.default_rng(12345)
    ntx = np.exp(rng.uniform(np.log(5000), np.log(200000), S)).astype(int)
    nfr = rng.integers(5_000_000, 80_000_001, S)
    gens = [H.generate(int(t), int(10 * t), int(f), seed=1000 + q)
            for q, (t, f) in enumerate(zip(ntx, nfr))]
    return [g[0] for g in gens], [g[1] for g in gens]


def test_ragged_64_samples_20_replicates():
    S, R = 64, 20
    cls, us = ragged_samples(S)
    say("ragged: %d samples, %d transcripts, %d classes, %d members" % (
        S, sum(c.n_tx for c in cls), sum(c.n_cls for c in cls),
        sum(c.nnz for c in cls)))
    seeds = (np.arange(S * R, dtype=np.uint64) + 1).reshape(S, R)
    mp = H.Problem.from_samples(cls)
    try:
        res = mp.bootstrap(np.concatenate(us), seeds, **STOP)
        # every sample's gradient in one call at its first replicate's
        # estimate (theta = log est, softmax within the sample)
        with np.errstate(divide="ignore"):
            th_all = np.log(res["est"][0])
        th_all = np.where(np.isfinite(th_all), th_all, -745.0)
        ell_all = np.log(np.concatenate([c.tx_len for c in cls]).astype(float))
        gs = mp.grad_samples(np.concatenate(us), th_all, ell_all)
    finally:
        mp.close()
    say("ragged gradient in one call: %.2f ms" % gs["ms"])
    it = res["iters"]
    say("ragged: iterations %d..%d, %.1f ms solve" % (
        it.min(), it.max(), res["stats"]["ms_solve"]))
    tx0 = np.concatenate([[0], np.cumsum([c.n_tx for c in cls])])
    # every (sample, replicate) instance equals the sample's own handle
    counts = {}
    sizes = np.array([c.n_tx for c in cls])
    check = [int(np.argmin(sizes)), int(np.argsort(sizes)[S // 2]),
             int(np.argmax(sizes))]
    for q in range(S):
        p = H.Problem(cls[q])
        try:
            own = p.bootstrap(us[q], seeds[q], **STOP)
            if q in check:
                counts[q] = p.bootstrap_counts(us[q], seeds[q][:2])
        finally:
            p.close()
        a, e = int(tx0[q]), int(tx0[q + 1])
        np.testing.assert_array_equal(it[q], own["iters"])
        np.testing.assert_array_equal(res["est"][:, a:e], own["est"])
        np.testing.assert_array_equal(res["tpm"][:, a:e], own["tpm"])
    # the smallest, median and largest sample: two replicates each
    # against the oracle on the same counts
    def oracle_pair(q):
        uo, co = counts[q]
        return O.bootstrap_batch(cls[q], uo, co, 2, **ORA_STOP)

    with ThreadPoolExecutor(len(check)) as ex:
        outs = list(ex.map(oracle_pair, check))
    for q, (est, ks) in zip(check, outs):
        a, e = int(tx0[q]), int(tx0[q + 1])
        say("sample %d (T=%d): oracle k %s, GPU %s" % (
            q, cls[q].n_tx, ks.tolist(), it[q, :2].tolist()))
        np.testing.assert_array_equal(ks, it[q, :2])
        np.testing.assert_array_equal(res["est"][:2, a:e], est)
    # the one-call gradient against each sample's own handle at the same
    # point; the median sample also against the CPU gradient
    for q in range(S):
        a, e = int(tx0[q]), int(tx0[q + 1])
        th, ell = th_all[a:e], ell_all[a:e]
        p = H.Problem(cls[q])
        try:
            g = p.grad(us[q], th, ell)
        finally:
            p.close()
        gt, ge = gs["g_theta"][a:e], gs["g_log_efflen"][a:e]
        assert np.all(np.isfinite(gt)) and np.all(np.isfinite(ge))
        # softmax invariance, and d/d(log efflen) = -E[count] <= 0
        assert abs(gt.sum()) <= 1e-9 * np.abs(gt).sum() + 1e-9
        assert np.all(ge <= 0)
        np.testing.assert_allclose(gt, g["g_theta"], rtol=0,
                                   atol=1e-10 * np.abs(g["g_theta"]).max())
        np.testing.assert_allclose(ge, g["g_log_efflen"], rtol=1e-12)
        assert gs["loglik"][q] == pytest.approx(g["loglik"], rel=1e-12)
        if q == check[1]:
            L, ogt, oge = O.grad(cls[q], us[q], th, ell)
            np.testing.assert_allclose(gt, ogt, rtol=1e-10,
                                       atol=1e-10 * np.abs(ogt).max())
            np.testing.assert_allclose(ge, oge, rtol=1e-10,
                                       atol=1e-10 * np.abs(oge).max())
            assert gs["loglik"][q] == pytest.approx(L, rel=1e-12)
