"""hem_bootstrap_hera on the GPU against the oracle's restatement of hera's
own sampler (src/EM.c:98-150, pinned to the C library's rand_r in
tests/test_hera_sampler.py): counts bit for bit, the seed chain of
replicates run one after another, class counts left alone, and every
replicate's EM bit for bit the oracle's on those counts.
"""
import sys

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed", [0, 987654321])
def test_hera_counts_bitwise(cfg1, seed):
    cls, u = cfg1
    p = H.Problem(cls)
    got = p.bootstrap_hera_counts(u, 3, seed)
    p.close()
    ref, _ = O.hera_sample(u, 3, seed)
    say("N = %d, first row sum %d" % (int(u.sum()), int(got[0].sum())))
    np.testing.assert_array_equal(got, ref)


def test_hera_seed_chain(cfg1):
    """Replicate r from seed equals replicate 0 from the state hera's
    global seed holds after r replicates."""
    cls, u = cfg1
    _, s1 = O.hera_sample(u, 1, 0)
    _, s2 = O.hera_sample(u, 2, 0)
    p = H.Problem(cls)
    a = p.bootstrap_hera_counts(u, 3, 0)
    b = p.bootstrap_hera_counts(u, 1, s1)
    c = p.bootstrap_hera_counts(u, 1, s2)
    p.close()
    np.testing.assert_array_equal(a[1], b[0])
    np.testing.assert_array_equal(a[2], c[0])


def test_hera_bootstrap_end_to_end_bitwise(cfg1):
    cls, u = cfg1
    p = H.Problem(cls)
    res = p.bootstrap_hera(u, 3, 0)
    p.close()
    ref, _ = O.hera_sample(u, 3, 0)
    say("iters %s resample %.2f ms" % (res["iters"].tolist(),
                                       res["stats"]["ms_resample"]))
    for r in range(3):
        ora = O.estimate_count(cls, ref[r])   # class counts unchanged
        assert res["iters"][r] == ora["k"]
        np.testing.assert_array_equal(res["est"][r], ora["est"])
        np.testing.assert_array_equal(res["tpm"][r], ora["tpm"])


def test_hera_small_and_rejects(cfg1):
    members = [[0, 1], [1, 2], [3, 4]]
    c = H.Classes.from_lists(5, members, [3.0, 2.0, 4.0], [100] * 5)
    p = H.Problem(c)
    u1 = np.array([0, 1, 0, 0, 0], np.uint32)   # one read: no draws
    assert (p.bootstrap_hera_counts(u1, 2, 5) == 0).all()
    u0 = np.zeros(5, np.uint32)
    assert (p.bootstrap_hera_counts(u0, 2, 5) == 0).all()
    u = np.array([4, 0, 9, 1, 2], np.uint32)
    got = p.bootstrap_hera_counts(u, 4, 11)
    np.testing.assert_array_equal(got, O.hera_sample(u, 4, 11)[0])
    p.close()
    cls, uu = cfg1
    mp = H.Problem.from_samples([cls, cls])
    with pytest.raises(H.HemError) as ei:
        mp.bootstrap_hera_counts(np.concatenate([uu, uu]), 2, 0)
    assert ei.value.code == _lib.HEM_ERR_INVALID
    mp.close()
