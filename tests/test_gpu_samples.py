"""Ragged multi-sample batches (hem_samples_create): several samples' class
tables side by side in one handle, every sample's replicates solved
together.  Each (sample, replicate) instance must be bitwise what the
sample's own handle computes for the same counts and seeds -- and that, in
the exact modes, is bitwise the oracle (tests/test_gpu_batch.py).
"""
import sys

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [(2000, 20000, 1_000_000, 3), (5000, 50000, 3_000_000, 4),
         (800, 8000, 500_000, 5)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def samples():
    out = [H.generate(t, c, f, seed=s) for t, c, f, s in SIZES]
    return [c for c, _ in out], [u for _, u in out]


def per_sample_seeds(n, R):
    return (np.arange(n * R, dtype=np.uint64).reshape(n, R) * 7919 + 11)


@pytest.mark.parametrize("norm", [H.NORM_TREE, H.NORM_EXACT])
def test_samples_bootstrap_equals_own_handles(samples, norm):
    cls, us = samples
    R = 4
    seeds = per_sample_seeds(len(cls), R)
    mp = H.Problem.from_samples(cls)
    u_all = np.concatenate(us)
    res = mp.bootstrap(u_all, seeds, norm=norm, stop_rel_change=1e-6)
    uo, co = mp.bootstrap_counts(u_all, seeds)
    mp.close()
    say("norm %d iters %s" % (norm, res["iters"].tolist()))
    for q, (c, u) in enumerate(zip(cls, us)):
        t0 = int(np.sum([x.n_tx for x in cls[:q]]))
        c0 = int(np.sum([x.n_cls for x in cls[:q]]))
        p = H.Problem(c)
        own = p.bootstrap(u, seeds[q], norm=norm, stop_rel_change=1e-6)
        ou, oc = p.bootstrap_counts(u, seeds[q])
        p.close()
        np.testing.assert_array_equal(uo[:, t0:t0 + c.n_tx], ou)
        np.testing.assert_array_equal(co[:, c0:c0 + c.n_cls], oc)
        np.testing.assert_array_equal(res["iters"][q], own["iters"])
        np.testing.assert_array_equal(res["est"][:, t0:t0 + c.n_tx],
                                      own["est"])
        np.testing.assert_array_equal(res["tpm"][:, t0:t0 + c.n_tx],
                                      own["tpm"])
        # tpm is normalised within each sample
        np.testing.assert_allclose(res["tpm"][:, t0:t0 + c.n_tx].sum(1),
                                   1e6, rtol=1e-9)


def test_samples_solve_batch_equals_oracle(samples):
    """Shared counts (each sample's own), one replicate: every sample's
    instance is the reference EM of that sample, bit for bit."""
    cls, us = samples
    mp = H.Problem.from_samples(cls)
    res = mp.solve_batch(np.concatenate(us)[None, :])
    mp.close()
    for q, (c, u) in enumerate(zip(cls, us)):
        t0 = int(np.sum([x.n_tx for x in cls[:q]]))
        ora = O.estimate_count(c, u)
        assert res["iters"][q, 0] == ora["k"]
        np.testing.assert_array_equal(res["est"][0, t0:t0 + c.n_tx],
                                      ora["est"])
        np.testing.assert_array_equal(res["tpm"][0, t0:t0 + c.n_tx],
                                      ora["tpm"])


@pytest.mark.parametrize("prec", [H.PREC_FP64, H.PREC_FP32])
def test_samples_gradient_equals_own_handles(samples, prec):
    """One pass over the multi-sample handle gives each sample the gradient
    its own handle gives (softmax within the sample), to rounding."""
    cls, us = samples
    rng = np.random.default_rng(21)
    ths = [rng.normal(0.0, 1.0, c.n_tx) for c in cls]
    ells = [np.log(c.tx_len.astype(float)) + rng.normal(0.0, 0.1, c.n_tx)
            for c in cls]
    mp = H.Problem.from_samples(cls)
    try:
        g = mp.grad_samples(np.concatenate(us), np.concatenate(ths),
                            np.concatenate(ells), precision=prec)
        with pytest.raises(H.HemError):    # hem_grad is one-sample only
            mp.grad(np.concatenate(us), np.concatenate(ths),
                    np.concatenate(ells))
    finally:
        mp.close()
    say("per-sample loglik %s, %.3f ms (adjoint %.3f ms)" % (
        g["loglik"].tolist(), g["ms"], g["ms_adjoint"]))
    t0 = 0
    for q, (c, u) in enumerate(zip(cls, us)):
        p = H.Problem(c)
        own = p.grad(u, ths[q], ells[q], precision=prec)
        p.close()
        sl = slice(t0, t0 + c.n_tx)
        t0 += c.n_tx
        assert g["loglik"][q] == pytest.approx(own["loglik"], rel=1e-12)
        assert g["n_frag"][q] == pytest.approx(own["n_frag"], rel=1e-12)
        gt = own["g_theta"]
        np.testing.assert_allclose(g["g_theta"][sl], gt, rtol=0,
                                   atol=1e-10 * np.abs(gt).max())
        np.testing.assert_allclose(g["g_log_efflen"][sl],
                                   own["g_log_efflen"], rtol=1e-12)
        # softmax invariance within each sample
        assert abs(g["g_theta"][sl].sum()) <= 1e-9 * np.abs(gt).sum()
    if prec == H.PREC_FP64:
        # the middle sample against the CPU gradient
        L, ogt, oge = O.grad(cls[1], us[1], ths[1], ells[1])
        sl = slice(cls[0].n_tx, cls[0].n_tx + cls[1].n_tx)
        assert g["loglik"][1] == pytest.approx(L, rel=1e-12)
        np.testing.assert_allclose(g["g_theta"][sl], ogt, rtol=0,
                                   atol=1e-10 * np.abs(ogt).max())


def test_samples_reject_single_problem_entry_points(samples):
    cls, us = samples
    mp = H.Problem.from_samples(cls)
    with pytest.raises(H.HemError) as ei:
        mp.solve(np.concatenate(us))
    assert ei.value.code == _lib.HEM_ERR_INVALID
    mp.close()
