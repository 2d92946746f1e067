"""The monotonicity fallback on the GPU (hem_opts.fallback = 1; no
counterpart in the reference, pinned in tests/test_oracle.py): every
iterate bitwise the oracle's with the same option, the same iterations
falling back, and the refusals where the option does not run.
"""
import sys

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from conftest import tiny_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve_fb(cls, u, k=None, **kw):
    """GPU and oracle with the fallback, on the oracle's schedule (k
    iterations) and under the stop rule."""
    okw = dict(kw)
    gkw = {"stop_rel": "stop_rel_change", "min_round": "min_iter",
           "max_round": "max_iter"}
    ora = O.estimate_count(cls, u, fallback=1, **okw)
    p = H.Problem(cls)
    try:
        g = p.solve(u, fallback=1, **{gkw[a]: b for a, b in okw.items()})
        prof = p.profile(u, fallback=1, **{gkw[a]: b for a, b in
                                           okw.items()})
    finally:
        p.close()
    assert prof["tile_ex3_map12"]["launches"] > 0
    say("T=%d k gpu %d oracle %d; fallbacks gpu %d oracle %d" % (
        cls.n_tx, g["stats"]["iters"], ora["k"], g["stats"]["fallbacks"],
        int(ora["fallback"].sum())))
    return g, ora


@pytest.mark.parametrize("seed", [13, 23, 30])
def test_fallback_bitwise_on_overshooting_tables(seed):
    cls, u = tiny_problem(seed, n_tx=60, n_cls=300, zero_frac=0.3)
    for kw in (dict(), dict(min_round=40, max_round=40)):
        g, ora = solve_fb(cls, u, **kw)
        assert ora["fallback"].sum() > 0
        assert g["stats"]["iters"] == ora["k"]
        assert g["stats"]["fallbacks"] == ora["fallback"].sum()
        np.testing.assert_array_equal(g["lam"], ora["lam"])
        np.testing.assert_array_equal(g["tpm"], ora["tpm"])
        np.testing.assert_array_equal(g["est"], ora["est"])


@pytest.mark.parametrize("stop_rel", [0.01, 1e-8])
def test_fallback_bitwise_cfg1(cfg1, stop_rel):
    cls, u = cfg1
    g, ora = solve_fb(cls, u, stop_rel=stop_rel)
    assert g["stats"]["iters"] == ora["k"]
    assert g["stats"]["fallbacks"] == ora["fallback"].sum()
    np.testing.assert_array_equal(g["lam"], ora["lam"])
    np.testing.assert_array_equal(g["est"], ora["est"])


def test_fallback_bitwise_weighted():
    cls, u = H.generate(20000, 200000, 10_000_000, seed=7)
    rng = np.random.default_rng(7)
    cls.member_weight = rng.uniform(0.05, 1.0, cls.nnz)
    g, ora = solve_fb(cls, u, stop_rel=1e-8)
    assert g["stats"]["iters"] == ora["k"]
    assert g["stats"]["fallbacks"] == ora["fallback"].sum()
    np.testing.assert_array_equal(g["lam"], ora["lam"])


def test_fallback_off_unchanged(cfg1):
    cls, u = cfg1
    p = H.Problem(cls)
    try:
        a = p.solve(u)
        b = p.solve(u, fallback=0)
        c = p.solve(u, fallback=1)
        d = p.solve(u)
    finally:
        p.close()
    np.testing.assert_array_equal(a["lam"], b["lam"])
    np.testing.assert_array_equal(a["lam"], d["lam"])
    assert a["stats"]["fallbacks"] == 0 and d["stats"]["fallbacks"] == 0
    assert c["stats"]["iters"] >= 50


def test_fallback_refusals():
    # a component over the tile limits (the global passes), FP32, and
    # batched solves do not take the option
    cls, u = tiny_problem(3, n_tx=200, n_cls=1500)
    p = H.Problem(cls)
    try:
        with pytest.raises(H.HemError) as ei:
            p.solve(u, fallback=1)
        assert ei.value.code == _lib.HEM_ERR_RANGE
    finally:
        p.close()
    cls, u = tiny_problem(1, n_tx=50, n_cls=200)
    p = H.Problem(cls)
    try:
        with pytest.raises(H.HemError) as ei:
            p.solve(u, fallback=1, precision=H.PREC_FP32)
        assert ei.value.code == _lib.HEM_ERR_RANGE
        with pytest.raises(H.HemError) as ei:
            p.solve(u, fallback=2)
        assert ei.value.code == _lib.HEM_ERR_INVALID
        with pytest.raises(H.HemError):
            p.solve_batch(np.stack([u, u]), fallback=1)
    finally:
        p.close()
