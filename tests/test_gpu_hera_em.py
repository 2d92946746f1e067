"""hera-EM's simplified-model SQUAREM with locking and read filtering on
the GPU (hem_solve_hera_em): every result bitwise the oracle's restatement
(pinned to hera-EM's own per-read loop in tests/test_hera_em_mode.py), with
the same loop count, locks and removals."""
import sys

import numpy as np
import pytest

import hera_amd as H
from conftest import hera_em_reads, hera_em_table, tiny_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GKEY = {"lock_thr": "lock_threshold", "filter_thr": "filter_threshold",
        "stop_rel": "stop_rel_change", "stop_min": "stop_min_value",
        "max_iter": "max_iter", "locking": "locking",
        "filtering": "filtering"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(cls, noise, p0=None, **kw):
    po, so = O.hera_em(cls, noise, p0=p0, **kw)
    p = H.Problem(cls)
    try:
        pg, sg = p.solve_hera_em(noise, p0=p0,
                                 **{GKEY[k]: v for k, v in kw.items()})
        # a second solve on the same handle starts afresh
        pg2, sg2 = p.solve_hera_em(noise, p0=p0,
                                   **{GKEY[k]: v for k, v in kw.items()})
    finally:
        p.close()
    say("T=%d C=%d iters %d/%d locked %d/%d removed members %d/%d "
        "classes %d/%d" % (cls.n_tx, cls.n_cls, sg["iters"], so["iters"],
                           sg["n_locked"], so["n_locked"],
                           sg["removed_members"], so["removed_members"],
                           sg["removed_classes"], so["removed_classes"]))
    for key in ("iters", "converged", "n_locked", "removed_members",
                "removed_classes"):
        assert sg[key] == so[key], key
        assert sg2[key] == so[key], key
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(pg2, po)
    return pg, sg


@pytest.mark.parametrize("seed", [3, 9, 21])
def test_hera_em_small_bitwise(seed):
    T = 30
    cls, _ = hera_em_table(T, hera_em_reads(seed, n_tx=T, n_pat=120,
                                           max_aln=4))
    _, st = both(cls, T)
    assert st["n_locked"] > 0 and st["removed_members"] > 0


def test_hera_em_large_bitwise():
    T = 500
    cls, _ = hera_em_table(T, hera_em_reads(9, n_tx=T, n_pat=4000,
                                           max_aln=6))
    _, st = both(cls, T)
    assert st["converged"] == 1


@pytest.mark.parametrize("kw", [dict(locking=0), dict(filtering=0),
                                dict(locking=0, filtering=0),
                                dict(filter_thr=1e-3, lock_thr=1e-3),
                                dict(stop_rel=0.0, max_iter=15)])
def test_hera_em_options_bitwise(kw):
    T = 60
    cls, _ = hera_em_table(T, hera_em_reads(5, n_tx=T, n_pat=400,
                                           max_aln=5))
    both(cls, T, **kw)


def test_hera_em_start_point_and_no_noise():
    T = 60
    cls, _ = hera_em_table(T, hera_em_reads(7, n_tx=T, n_pat=300,
                                           max_aln=5))
    rng = np.random.default_rng(1)
    p0 = rng.uniform(0.1, 1.0, T + 1)
    p0 /= p0.sum()
    both(cls, T, p0=p0)
    # an unweighted table without a noise component
    cls2, _ = tiny_problem(4, n_tx=40, n_cls=150)
    both(cls2, cls2.n_tx)
