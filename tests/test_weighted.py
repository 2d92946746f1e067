"""Class tables with per-member weights, in the oracle (host only).

The weighted map gives member k of class c the share lam_t w_k n_c / D_c,
D_c = sum_k lam_{t_k} w_k.  Pinned here by: w = 1 reproducing the
unweighted estimate_count bit for bit; one map equal, after
normalisation, to a restatement of hera-EM's simple-model E-step with noise
(hera-EM/em.c:75-157, smodel.h:15-45, transdis.c:54-61); and the SQUAREM
solve reaching the fixed point that hera-EM's plain EM iterates to.
"""
import numpy as np

from conftest import hera_em_reads, hera_em_table, tiny_problem
from oracle import oracle as O

N_TX = 30


def hera_em_step(g, gn, pats):
    """hera-EM's e_step without filtering or locking: per read, total =
    sum_k g_t c_k + g_noise 4^-len; each alignment adds g_t c_k / total, the
    noise g_noise 4^-len / total; then every entry, noise included, is
    divided by their sum."""
    new = np.zeros(len(g))
    new_n = 0.0
    for aln, rl, n in pats:
        probs = [g[t] * c for t, c in aln]
        total = 0.0
        for p in probs:
            total += p
        noise = gn * 0.25 ** rl
        total += noise
        for _ in range(n):
            new_n += noise / total
            if total < 1e-300:
                continue
            for (t, _), p in zip(aln, probs):
                new[t] += p / total
    s = new.sum() + new_n
    return new / s, new_n / s


def test_unit_weights_are_hera_bitwise():
    cls, u = tiny_problem(11, n_tx=60, n_cls=300, dup=True)
    ref = O.estimate_count(cls, u)
    cls.member_weight = np.ones(cls.nnz)
    got = O.estimate_count(cls, u)
    assert got["k"] == ref["k"]
    np.testing.assert_array_equal(got["lam"], ref["lam"])
    np.testing.assert_array_equal(got["tpm"], ref["tpm"])


def test_weighted_map_is_hera_em_estep():
    pats = hera_em_reads(1, N_TX)
    cls, u = hera_em_table(N_TX, pats)
    rng = np.random.default_rng(2)
    lam = rng.uniform(0.1, 1.0, N_TX + 1)
    out, _ = O.em_map(cls, u, lam)
    g = lam / lam.sum()
    want, want_n = hera_em_step(g[:N_TX], g[N_TX], pats)
    got = out / out.sum()
    np.testing.assert_allclose(got[:N_TX], want, rtol=1e-12, atol=0)
    np.testing.assert_allclose(got[N_TX], want_n, rtol=1e-12)
    # the map returns hera-EM's unnormalised counts: sum = number of reads
    n_reads = sum(n for _, _, n in pats)
    assert abs(out.sum() - n_reads) < 1e-9 * n_reads


def test_weighted_map_changes_with_weights():
    """A dropped weight would leave the map equal to the unweighted one."""
    pats = hera_em_reads(3, N_TX)
    cls, u = hera_em_table(N_TX, pats)
    lam = np.full(N_TX + 1, 0.5)
    a, _ = O.em_map(cls, u, lam)
    cls.member_weight = np.ones(cls.nnz)
    b, _ = O.em_map(cls, u, lam)
    assert np.abs(a - b).max() > 1e-3


def test_weighted_solve_reaches_hera_em_fixed_point():
    """In hera-EM's own SQUAREM forms (L2 step length, u2/10 in place of a
    non-positive extrapolate, no step clamp: hera-EM/utility.c:16-53) the
    weighted solve ends at the fixed point of hera-EM's plain EM.  (hera's
    clamp at zero can park a transcript at 0, which the map keeps.)"""
    pats = hera_em_reads(5, N_TX, n_pat=120)
    cls, u = hera_em_table(N_TX, pats)
    res = O.estimate_count(cls, u, stop_min=0.0, stop_rel=1e-12,
                           max_round=1000, step_form=1, clamp_form=1,
                           alpha_clamp=0)
    got = res["lam"] / res["lam"].sum()
    g = np.full(N_TX, 1.0 / (N_TX + 1))
    gn = 1.0 / (N_TX + 1)
    for _ in range(3000):
        g, gn = hera_em_step(g, gn, pats)
    big = g > 1e-3
    np.testing.assert_allclose(got[:N_TX][big], g[big], rtol=1e-6)
    np.testing.assert_allclose(got[N_TX], gn, rtol=1e-6)
