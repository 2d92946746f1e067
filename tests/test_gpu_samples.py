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


def test_samples_reject_single_problem_entry_points(samples):
    cls, us = samples
    mp = H.Problem.from_samples(cls)
    with pytest.raises(H.HemError) as ei:
        mp.solve(np.concatenate(us))
    assert ei.value.code == _lib.HEM_ERR_INVALID
    mp.close()
