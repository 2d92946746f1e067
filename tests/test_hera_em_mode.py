"""hera-EM's simplified-model SQUAREM with locking and read filtering,
restated in the oracle (oracle_hera_em, hera-EM/em.c:75-157, 384-426;
smodel.h:15-45; transdis.c:20-61, 97-130; utility.c:64-128; utility.h:
10-14, 86-122).

The oracle works on classes (identical reads collapsed into a count).  It
is pinned here against an independent transcription of hera-EM's own loop
that keeps every read separately, with hera-EM's alignment lists that
filtering shortens in place, per-read locked mass (read->prob) and the
banked noise count (noise_cnt): the two must agree to rounding (the
transcription adds a class's reads one at a time, the oracle once with
the count) on iteration count, locks, removals and the final
distribution; and by invariants: the distribution sums to 1, locked
entries keep their value, filtering only removes what its rules name, and
with locking and filtering off and the threshold at 0 the fixed point is
the one of the plain weighted map.
"""
import numpy as np
import pytest

from conftest import hera_em_reads, hera_em_table
from oracle import oracle as O

EPS = 1e-300


def he_reads(pats):
    """hera-EM's read list: every read of a pattern on its own, in pattern
    order, with its alignments (t, c_prob), noise weight 4^-len and locked
    mass 0."""
    reads = []
    for aln, rl, n in pats:
        for _ in range(n):
            reads.append({"aln": list(aln), "nw": 0.25 ** rl, "prob": 0.0})
    return reads


def he_estep(reads, state, p, lock, filt, thr, z):
    """The e_step macro of hera-EM/em.c:75-157 for the simple model."""
    new = np.zeros_like(p)
    # n_update(new_m, NULL, noise_cnt): s_model_update_noise skips a count
    # below the threshold and a locked noise
    nc = state["noise_cnt"]
    if not nc < thr and not lock[z]:
        new[z] += nc
    kept = []
    for read in reads:
        total = read["prob"]
        prob = []
        for t, w in read["aln"]:
            x = p[t] * w
            prob.append(x)
            total += x
        noise = p[z] * read["nw"]
        total += noise
        if not noise / total < thr and not lock[z]:
            new[z] += noise / total
        if total < EPS:
            state["noise_cnt"] += filt
            kept.append(read)
            continue
        cnt = 0
        keep = []
        for (t, w), x in zip(read["aln"], prob):
            post = x / total
            mode = 2 if post < thr else (1 if lock[t] else 0)
            if mode == 0:
                new[t] += post
            if filt:
                if mode == 0:
                    keep.append((t, w))
                    cnt += 1
                elif mode == 1:
                    read["prob"] += x
                    state["removed_members"] += 1
                else:
                    state["removed_members"] += 1
        if filt:
            if cnt == 0:
                state["noise_cnt"] += noise / total
                state["removed_reads"] += 1
            else:
                read["aln"] = keep
                kept.append(read)
        else:
            kept.append(read)
    reads[:] = kept
    return new


def he_mstep(p, new, lock):
    """trans_master_init + lock_normalize: locked entries keep p."""
    out = np.where(lock, p, new)
    s = 0.0
    ls = 0.0
    for i in range(len(p)):
        if lock[i]:
            ls += out[i]
        else:
            s += out[i]
    f = 1 / s * (1 - ls)
    return np.where(lock, out, out * f)


def he_run(pats, T, z, max_iter=1000, locking=True, filtering=True,
           thr=1e-6, lock_thr=1e-5, t1=1e-3, t2=1e-7):
    reads = he_reads(pats)
    state = {"noise_cnt": 0.0, "removed_members": 0, "removed_reads": 0}
    lock = np.zeros(T, bool)
    p1 = np.full(T, 1.0 / T)
    k = 0
    conv = False
    while k < max_iter:
        p2 = he_mstep(p1, he_estep(reads, state, p1, lock, 0, thr, z), lock)
        p3 = he_mstep(p2, he_estep(reads, state, p2, lock, int(filtering),
                                   thr, z), lock)
        r2 = v2 = 0.0
        for i in range(T):
            if not lock[i]:
                r = p2[i] - p1[i]
                v = p3[i] - p2[i] - r
                r2 += r * r
                v2 += v * v
        g2 = 1.0 if v2 <= EPS else r2 / v2
        g = -np.sqrt(g2)
        for i in range(T):
            if not lock[i]:
                r = p2[i] - p1[i]
                v = p3[i] - p2[i] - r
                nv = p1[i] - 2 * g * r + g2 * v
                p3[i] = p3[i] / 10 if nv <= EPS else nv
        if locking:
            for i in range(T):
                if not lock[i]:
                    lock[i] = p3[i] < EPS or abs(p3[i] - p1[i]) < \
                        lock_thr * p3[i]
        p1 = he_mstep(p3, he_estep(reads, state, p3, lock, 0, thr, z), lock)
        k += 1
        bad = 0
        for i in range(T):
            if p3[i] >= t2:
                bad += abs(p3[i] - p1[i]) / p3[i] >= t1
        if bad == 0:
            conv = True
            break
    return p1, dict(iters=k, converged=int(conv), n_locked=int(lock.sum()),
                    removed_members=state["removed_members"])


@pytest.mark.parametrize("seed", [3, 9, 21])
def test_oracle_matches_per_read_transcription(seed):
    T = 30
    pats = hera_em_reads(seed, n_tx=T, n_pat=120, max_aln=4)
    cls, _ = hera_em_table(T, pats)
    p_py, st_py = he_run(pats, T + 1, T)
    p_or, st_or = O.hera_em(cls, T)
    assert st_or["iters"] == st_py["iters"]
    assert st_or["converged"] == st_py["converged"] == 1
    assert st_or["n_locked"] == st_py["n_locked"] > 0
    # a removed member of a class is removed for each of its reads
    assert st_or["removed_members"] > 0
    np.testing.assert_allclose(p_or, p_py, rtol=1e-11, atol=1e-15)


def test_filtering_and_locking_change_the_result():
    T = 30
    pats = hera_em_reads(9, n_tx=T, n_pat=120, max_aln=4)
    cls, _ = hera_em_table(T, pats)
    p, st = O.hera_em(cls, T)
    p0, st0 = O.hera_em(cls, T, locking=0, filtering=0)
    assert st0["n_locked"] == 0 and st0["removed_members"] == 0
    assert st["n_locked"] > 0 and st["removed_members"] > 0
    assert not np.array_equal(p, p0)
    # both are distributions
    assert p.sum() == pytest.approx(1.0, abs=1e-12)
    assert p0.sum() == pytest.approx(1.0, abs=1e-12)
    # the iterations they take are stopped by hera-EM's rule
    assert st["converged"] == 1 and st0["converged"] == 1


def test_without_locking_or_filtering_the_fixed_point_is_the_weighted_maps():
    # threshold 0, no locks, no filtering: hera-EM's loop is SQUAREM on the
    # weighted map (normalised after each map); its fixed point is the one
    # plain EM reaches (20,000 normalised weighted maps)
    T = 30
    pats = hera_em_reads(21, n_tx=T, n_pat=150, max_aln=4)
    cls, u = hera_em_table(T, pats)
    p, st = O.hera_em(cls, T, locking=0, filtering=0, filter_thr=0.0,
                      stop_rel=1e-13, stop_min=0.0, max_iter=20000)
    assert st["converged"] == 1
    lam = np.full(T + 1, 1.0 / (T + 1))
    for _ in range(20000):
        lam, _ = O.em_map(cls, u, lam)
        lam = lam / lam.sum()
    np.testing.assert_allclose(p, lam, rtol=0, atol=1e-12)


def test_stop_rule_is_arr_diff():
    # is_stop (transdis.c:123-130): no entry with p3 >= 1e-7 may change by
    # >= 1e-3 relative to p3; the restated loop stops only then, and a
    # stop_rel of 0 runs to the iteration cap
    T = 30
    pats = hera_em_reads(3, n_tx=T, n_pat=120, max_aln=4)
    cls, _ = hera_em_table(T, pats)
    _, st = O.hera_em(cls, T, stop_rel=0.0, max_iter=25)
    assert st["iters"] == 25 and st["converged"] == 0
    _, st = O.hera_em(cls, T, stop_rel=1e-3, max_iter=1000)
    _, st2 = O.hera_em(cls, T, stop_rel=1e-6, max_iter=1000)
    assert st["converged"] == st2["converged"] == 1
    assert st2["iters"] > st["iters"]


def test_locked_entries_keep_their_value():
    # an entry locked by the first loop's extrapolation keeps its value
    # through every later map: with a huge lock threshold everything locks
    # after one loop, and the distribution then never moves
    T = 30
    pats = hera_em_reads(3, n_tx=T, n_pat=120, max_aln=4)
    cls, _ = hera_em_table(T, pats)
    p1, st1 = O.hera_em(cls, T, lock_thr=1e9, max_iter=1, stop_rel=0.0)
    p5, st5 = O.hera_em(cls, T, lock_thr=1e9, max_iter=5, stop_rel=0.0)
    assert st1["n_locked"] == T + 1
    np.testing.assert_array_equal(p1, p5)
