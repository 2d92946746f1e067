"""Class tables with per-member weights on the GPU (hem_classes
member_weight): the FP64 exact solve bit for bit the oracle's weighted
estimate_count, unit weights bit for bit the unweighted solve, and the
entry points that do not take weights refusing them.  The oracle's weighted
map is pinned to hera-EM's simple-model E-step in tests/test_weighted.py.
"""
import sys

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from conftest import hera_em_reads, hera_em_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# hera's SQUAREM forms, and hera-EM's (L2 step, u2/10 clamp, no step clamp)
FORMS = [dict(), dict(step_form=1, clamp_form=1, alpha_clamp=0)]
GKEY = dict(step_form="step", clamp_form="clamp", alpha_clamp="alpha_clamp",
            min_round="min_iter", max_round="max_iter", stop_min="stop_min_value",
            stop_rel="stop_rel_change")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve_both(cls, u, tiled=True, **kw):
    ora = O.estimate_count(cls, u, **kw)
    p = H.Problem(cls)
    assert p.tiles()["n_tiles"] == 0      # no batched tiling for weights
    gkw = {GKEY[k]: v for k, v in kw.items()}
    g = p.solve(u, **gkw)
    # which kernels the solve ran: the tile kernels (weights in the tile
    # records) or the two global passes
    prof = p.profile(u, **gkw)
    p.close()
    assert (prof["tile_ex3_map12"]["launches"] > 0) == tiled
    assert (prof["class_pass"]["launches"] > 0) == (not tiled)
    say("T = %d, nnz = %d, k = %d / %d, tiled %s" % (
        cls.n_tx, cls.nnz, g["stats"]["iters"], ora["k"], tiled))
    return g, ora


def weighted_cfg(seed):
    cls, u = H.generate(20000, 200000, 10_000_000, seed=seed)
    rng = np.random.default_rng(seed)
    cls.member_weight = rng.uniform(0.05, 1.0, cls.nnz)
    return cls, u


@pytest.mark.parametrize("kw", FORMS)
def test_weighted_generated_bitwise(kw):
    cls, u = weighted_cfg(7)
    g, ora = solve_both(cls, u, **kw)
    assert g["stats"]["iters"] == ora["k"]
    np.testing.assert_array_equal(g["lam"], ora["lam"])
    np.testing.assert_array_equal(g["tpm"], ora["tpm"])
    np.testing.assert_array_equal(g["est"], ora["est"])
    assert np.isnan(g["stats"]["loglik"])


@pytest.mark.parametrize("kw", FORMS)
def test_weighted_hera_em_table_bitwise(kw):
    pats = hera_em_reads(9, 500, n_pat=4000, max_aln=6)
    cls, u = hera_em_table(500, pats)
    kw = dict(kw, stop_min=0.0, stop_rel=1e-10)
    # the noise transcript is a member of every class: one component,
    # over the tile limits, so the global passes run
    g, ora = solve_both(cls, u, tiled=False, **kw)
    assert g["stats"]["iters"] == ora["k"]
    np.testing.assert_array_equal(g["lam"], ora["lam"])
    np.testing.assert_array_equal(g["tpm"], ora["tpm"])


def test_weighted_untiled_bitwise():
    """A component over the tile limits keeps the table on the global
    passes, whose class and transcript kernels carry the weights too."""
    cls, u = H.generate(2000, 20000, 1_000_000, seed=3)
    members = [cls.cls_ref[cls.cls_off[i]:cls.cls_off[i + 1]].tolist()
               for i in range(cls.n_cls)]
    members += [[i, i + 1] for i in range(150)]
    counts = np.concatenate([cls.cls_count, np.full(150, 3.0)])
    big = H.Classes.from_lists(cls.n_tx, members, counts.tolist(),
                               cls.tx_len.tolist())
    big.member_weight = np.random.default_rng(3).uniform(0.05, 1.0, big.nnz)
    g, ora = solve_both(big, u, tiled=False, min_round=200, max_round=200)
    assert g["stats"]["iters"] == ora["k"] == 200
    np.testing.assert_array_equal(g["lam"], ora["lam"])


def test_unit_weights_bitwise_unweighted():
    cls, u = H.generate(20000, 200000, 10_000_000, seed=0)
    p = H.Problem(cls)
    a = p.solve(u)
    p.close()
    cls.member_weight = np.ones(cls.nnz)
    p = H.Problem(cls)
    b = p.solve(u)
    p.close()
    assert a["stats"]["iters"] == b["stats"]["iters"]
    np.testing.assert_array_equal(a["lam"], b["lam"])
    np.testing.assert_array_equal(a["tpm"], b["tpm"])


def test_weighted_tree_norm_close():
    cls, u = weighted_cfg(8)
    ora = O.estimate_count(cls, u)
    p = H.Problem(cls)
    g = p.solve(u, norm=H.NORM_TREE)
    p.close()
    big = ora["tpm"] > 1e-2
    np.testing.assert_allclose(g["tpm"][big], ora["tpm"][big], rtol=1e-4)


def test_weighted_rejects():
    cls, u = weighted_cfg(9)
    p = H.Problem(cls)
    for call in (lambda: p.solve(u, precision=H.PREC_FP32),
                 lambda: p.loglik(u, np.ones(cls.n_tx)),
                 lambda: p.grad(u, np.zeros(cls.n_tx), np.zeros(cls.n_tx)),
                 lambda: p.solve_batch(np.stack([u, u])),
                 lambda: p.bootstrap_hera(u, 2, 0)):
        with pytest.raises(H.HemError) as ei:
            call()
        assert ei.value.code == _lib.HEM_ERR_INVALID
    p.close()
    with pytest.raises(H.HemError) as ei:
        H.Problem.from_samples([cls, cls])
    assert ei.value.code == _lib.HEM_ERR_INVALID
    bad = weighted_cfg(9)[0]
    bad.member_weight[5] = -1.0
    with pytest.raises(H.HemError) as ei:
        H.Problem(bad)
    assert ei.value.code == _lib.HEM_ERR_INVALID
    bad.member_weight[5] = np.nan
    with pytest.raises(H.HemError) as ei:
        H.Problem(bad)
    assert ei.value.code == _lib.HEM_ERR_INVALID


def test_zero_weights_bitwise():
    """Weights may be 0: a class whose members all weigh 0 has D_c == 0
    while its transcripts' densities are positive; the reference's map
    skips it (src/EM.c:279-280), which the tile kernels reproduce with
    an exact zero term."""
    cls, u = weighted_cfg(12)
    rng = np.random.default_rng(12)
    w = cls.member_weight
    w[rng.random(cls.nnz) < 0.1] = 0.0
    c0 = int(np.argmax(np.diff(cls.cls_off) >= 3))
    w[cls.cls_off[c0]:cls.cls_off[c0 + 1]] = 0.0
    g, ora = solve_both(cls, u)
    assert g["stats"]["iters"] == ora["k"]
    np.testing.assert_array_equal(g["lam"], ora["lam"])
