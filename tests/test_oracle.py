"""Pins the CPU oracle (oracle/em_oracle.c) without running the reference.

The compiled reference (src/EM.c) cannot be built in this image (it needs
HDF5 headers), and the reference ships no tests or vectors.  Each check below
is chosen so that a plausible restatement mistake fails it: a wrong share in
the class loop, a dropped unique count, a transposed r/v, a wrong norm, a
wrong clamp, a wrong stop-test input or loop bound.
"""
import numpy as np
import pytest
from scipy.optimize import brentq, minimize

from conftest import tiny_problem
from hera_amd import Classes
from oracle import oracle as O


def class_terms(cls, lam):
    """D_c and per-transcript sum_{c ni t} n_c / D_c (numpy, vectorised)."""
    off = cls.cls_off.astype(np.int64)
    ref = cls.cls_ref.astype(np.int64)
    sizes = np.diff(off)
    row = np.repeat(np.arange(cls.n_cls), sizes)
    D = np.bincount(row, weights=lam[ref], minlength=cls.n_cls)
    rho = np.where(D > 0, cls.cls_count / np.where(D > 0, D, 1), 0.0)
    s = np.bincount(ref, weights=rho[row], minlength=cls.n_tx)
    return D, s


def test_norm_is_sample_sd():
    rng = np.random.default_rng(1)
    for n in (2, 3, 17, 1000, 100000):
        x = rng.standard_normal(n) * 10 ** rng.uniform(-8, 3) + rng.uniform(-5, 5)
        assert O.norm_sd(x) == pytest.approx(np.std(x, ddof=1), rel=1e-12)
    # n-1 divisor, not n: two points {0, 2} have sample sd sqrt(2)
    assert O.norm_sd([0.0, 2.0]) == pytest.approx(np.sqrt(2.0), rel=1e-15)


def test_first_map_from_uniform_start_closed_form():
    # From lam = 1/T every member of class c gets n_c * mult / |c|, where
    # mult counts repeated members; lam' = (u + that) / len.
    for seed in range(3):
        cls, u = tiny_problem(seed, dup=True)
        T = cls.n_tx
        lam0 = np.full(T, 1.0 / T)
        out, _ = O.em_map(cls, u, lam0)
        share = np.zeros(T)
        for c in range(cls.n_cls):
            m = cls.cls_ref[cls.cls_off[c]:cls.cls_off[c + 1]]
            for t in m:
                share[t] += cls.cls_count[c] / len(m)
        np.testing.assert_allclose(out, (u + share) / cls.tx_len, rtol=1e-13)


def test_singleton_classes_one_map_fixed_point():
    # A single-member class hands its whole count to its member:
    # lam' = (u + n_c) / len regardless of lam (src/EM.c:276-290).
    rng = np.random.default_rng(5)
    T = 30
    members = [[t] for t in range(T)]
    counts = rng.integers(1, 100, T).astype(float)
    lens = rng.integers(200, 3000, T)
    cls = Classes.from_lists(T, members, counts, lens)
    u = rng.integers(0, 50, T).astype(np.uint32)
    lam = rng.uniform(0.1, 2.0, T)
    out, _ = O.em_map(cls, u, lam)
    # (lam * n) / lam rounds to n within an ulp
    np.testing.assert_allclose(out, (u + counts) / lens, rtol=4e-16)


def test_map_conserves_fragments_and_skips_zero_denominators():
    for seed in range(4):
        cls, u = tiny_problem(seed, dup=(seed % 2 == 1))
        rng = np.random.default_rng(seed)
        lam = rng.uniform(0, 1, cls.n_tx)
        lam[rng.random(cls.n_tx) < 0.3] = 0.0   # some D_c become 0
        out, _ = O.em_map(cls, u, lam)
        D, _ = class_terms(cls, lam)
        expect = u.sum() + cls.cls_count[D > 0].sum()
        assert np.dot(cls.tx_len.astype(float), out) == pytest.approx(
            expect, rel=1e-12)
        assert (D == 0).any()


def test_two_transcript_fixed_point():
    # u1, u2 unique reads, one shared class n12: the fixed point solves
    # L1 l1 + L2 l2 = u1 + u2 + n12 and l1 (L1 - n12 / (l1 + l2)) = u1.
    u1, u2, n12, L1, L2 = 37, 11, 250.0, 1500, 800
    S = u1 + u2 + n12

    def g(l1):
        l2 = (S - L1 * l1) / L2
        return l1 * L1 - l1 * n12 / (l1 + l2) - u1

    l1 = brentq(g, 1e-15, S / L1, xtol=1e-300, rtol=1e-15)
    l2 = (S - L1 * l1) / L2
    cls = Classes.from_lists(2, [[0, 1]], [n12], [L1, L2])
    u = np.array([u1, u2], dtype=np.uint32)
    r = O.estimate_count(cls, u, min_round=1, max_round=100000, stop_min=0.0,
                         stop_rel=1e-15)
    np.testing.assert_allclose(r["lam"], [l1, l2], rtol=1e-11)


def loglik_np(cls, u, lam):
    D, _ = class_terms(cls, lam)
    live = (cls.cls_count > 0) & (D > 0)
    a = np.sum(cls.cls_count[live] * np.log(D[live]))
    b = np.sum(u[u > 0] * np.log(lam[u > 0]))
    return a + b - np.dot(cls.tx_len.astype(float), lam)


def test_loglik_definition_and_em_monotone():
    # EM never lowers the objective whose EM map src/EM.c:271-302 is.
    for seed in range(4):
        cls, u = tiny_problem(seed, dup=(seed == 3))
        rng = np.random.default_rng(seed + 10)
        lam = rng.uniform(1e-3, 1.0, cls.n_tx)
        ll0 = O.loglik(cls, u, lam)
        assert ll0 == pytest.approx(loglik_np(cls, u, lam), rel=1e-13)
        for _ in range(5):
            lam, _ = O.em_map(cls, u, lam)
            ll1 = O.loglik(cls, u, lam)
            assert ll1 >= ll0 - 1e-9 * abs(ll0)
            ll0 = ll1


def test_converged_point_is_the_maximum():
    # KKT of the Poisson objective at the tightly converged point, and no
    # better point found by an independent optimiser.
    cls, u = tiny_problem(7, n_tx=8, n_cls=20, max_size=4)
    r = O.estimate_count(cls, u, min_round=1, max_round=200000, stop_min=0.0,
                         stop_rel=1e-14)
    lam = r["lam"]
    _, s = class_terms(cls, lam)
    grad = np.where(u > 0, u / np.maximum(lam, 1e-300), 0.0) + s - cls.tx_len
    supp = lam > 1e-12
    assert np.all(np.abs(grad[supp]) < 1e-6 * cls.tx_len[supp])
    assert np.all(grad[~supp] < 1e-6 * cls.tx_len[~supp])
    ll = O.loglik(cls, u, lam)

    def negll(x):
        return -loglik_np(cls, u, np.exp(x))

    x0 = np.log(np.full(cls.n_tx, (u.sum() + cls.cls_count.sum())
                        / cls.tx_len.sum()))
    best = minimize(negll, x0, method="BFGS", options={"gtol": 1e-10,
                                                       "maxiter": 20000})
    assert ll >= -best.fun - 1e-9 * abs(ll)


def test_alpha_minus_one_when_em_is_exact():
    # Singleton classes: u1 = u2 = fixed point, v = -r, so sd(v) = sd(r),
    # alpha = -1, and the loop stops at min_round with lam = (u + n)/len.
    rng = np.random.default_rng(3)
    T = 25
    cls = Classes.from_lists(T, [[t] for t in range(T)],
                             rng.integers(1, 90, T).astype(float),
                             rng.integers(200, 3000, T))
    u = rng.integers(1, 40, T).astype(np.uint32)
    r = O.estimate_count(cls, u)
    assert r["k"] == 50
    assert np.all(r["alpha"] == -1.0)
    np.testing.assert_allclose(r["lam"], (u + cls.cls_count) / cls.tx_len,
                               rtol=4e-16)


def test_squarem_step_uses_sample_sd():
    # One SQUAREM iteration from the uniform start (src/EM.c:330-352):
    # r = u1 - u, v = (u2 - u1) - r from two plain maps, the step length
    # -Norm(r)/Norm(v) with Norm the sample sd of src/hash_align.c:150-164
    # (mean-centred, n-1), clamped to <= -1, then u - 2 a r + a^2 v >= 0.
    # The table is chosen so that r and v have means far from zero: there
    # the L2 form -||r||/||v|| (hera-EM/utility.c:16-41) gives another
    # step, so an oracle that used it by default would fail here.
    cls, u = tiny_problem(0)
    T = cls.n_tx
    u0 = np.full(T, 1.0 / T)
    u1, _ = O.em_map(cls, u, u0)
    u2, _ = O.em_map(cls, u, u1)
    r = u1 - u0
    v = (u2 - u1) - r
    a_sd = -np.std(r, ddof=1) / np.std(v, ddof=1)
    a_l2 = -np.sqrt(np.sum(r * r)) / np.sqrt(np.sum(v * v))
    # preconditions: both forms escape the clamp and differ by far more
    # than rounding
    assert a_sd < -1 and a_l2 < -1
    assert abs(a_sd - a_l2) > 0.1 * abs(a_sd)
    res = O.estimate_count(cls, u, min_round=1, max_round=1)
    assert res["k"] == 1
    assert res["alpha"][0] == pytest.approx(min(-1.0, a_sd), rel=1e-12)
    a = float(res["alpha"][0])
    # the extrapolated point (the third map's input), in the reference's
    # expression order: no FMA on x86-64 without -mfma, so bit for bit
    ext = np.maximum(u0 - 2 * a * r + a * a * v, 0.0)
    np.testing.assert_array_equal(res["last_u"], ext)
    lam, _ = O.em_map(cls, u, ext)
    np.testing.assert_array_equal(res["lam"], lam)
    # the L2 option is hera-EM's form, and is not the default
    res2 = O.estimate_count(cls, u, min_round=1, max_round=1, step_form=1)
    assert res2["alpha"][0] == pytest.approx(min(-1.0, a_l2), rel=1e-12)
    assert abs(res2["alpha"][0] - res["alpha"][0]) > 0.1 * abs(a_sd)


def test_alpha_clamp_and_extrapolation_clamp():
    cls, u = tiny_problem(11, n_tx=60, n_cls=300, zero_frac=0.6)
    r = O.estimate_count(cls, u, min_round=30, max_round=30)
    assert r["k"] == 30
    assert np.all(r["alpha"] <= -1.0)
    assert np.any(r["alpha"] < -1.0)
    r2 = O.estimate_count(cls, u, min_round=30, max_round=30, alpha_clamp=0)
    assert np.any(r2["alpha"] > -1.0)
    # zero clamp: once an extrapolated value is clamped to 0 and the
    # transcript has no unique reads it stays exactly 0
    assert np.any((r["lam"] == 0) & (u == 0))
    assert np.all(r["last_u"] >= 0)
    # hera-EM's u2/10 replacement never leaves exact zeros
    r3 = O.estimate_count(cls, u, min_round=30, max_round=30, clamp_form=1)
    assert np.all(r3["lam"] > 0)


def stop_flag(lam_new, lam_in, smin=1e-3, srel=0.01):
    m = lam_new > smin
    return int(not np.any(np.abs(lam_new[m] - lam_in[m]) / lam_new[m] > srel))


def test_stop_rule_inputs_and_loop_bounds(cfg1):
    cls, u = cfg1
    r = O.estimate_count(cls, u)
    k = r["k"]
    assert 50 <= k < 1000
    # the flag that ended the loop is the third map's test between the
    # extrapolated point and its image (src/EM.c:292-297, 355)
    assert r["stop"][k - 1] == 1
    assert stop_flag(r["lam"], r["last_u"]) == 1
    assert np.all(r["stop"][49:k - 1] == 0)
    # fixed iteration count and no-convergence paths
    r5 = O.estimate_count(cls, u, min_round=3, max_round=3)
    assert r5["k"] == 3
    assert stop_flag(r5["lam"], r5["last_u"]) == r5["stop"][2]
    rn = O.estimate_count(cls, u, min_round=0, max_round=7, stop_rel=0.0)
    assert rn["k"] == 7 and rn["stop"][6] == 0
    r0 = O.estimate_count(cls, u, min_round=0, max_round=0)
    assert r0["k"] == 0
    np.testing.assert_array_equal(r0["lam"], np.full(cls.n_tx, 1 / cls.n_tx))


def test_finalisation():
    cls, u = tiny_problem(2)
    r = O.estimate_count(cls, u)
    assert r["tpm"].sum() == pytest.approx(1e6, rel=1e-12)
    np.testing.assert_array_equal(r["est"], r["lam"] * cls.tx_len)
    np.testing.assert_allclose(r["tpm"], r["lam"] / r["lam"].sum() * 1e6,
                               rtol=1e-14)
    # all-zero input: the sum==0 branch zeroes both outputs
    z = O.estimate_count(Classes.from_lists(3, [[0, 1]], [0.0], [100, 100, 100]),
                         np.zeros(3, np.uint32), min_round=2, max_round=2)
    assert np.all(z["tpm"] == 0) and np.all(z["est"] == 0)


def test_gradient_matches_central_differences():
    cls, u = tiny_problem(4, n_tx=30, n_cls=90)
    rng = np.random.default_rng(0)
    theta = rng.standard_normal(cls.n_tx)
    ell = np.log(cls.tx_len.astype(float))
    L, gt, ge = O.grad(cls, u, theta, ell)
    h = 1e-6
    for j in range(cls.n_tx):
        e = np.zeros(cls.n_tx)
        e[j] = h
        fd_t = (O.grad(cls, u, theta + e, ell)[0]
                - O.grad(cls, u, theta - e, ell)[0]) / (2 * h)
        fd_e = (O.grad(cls, u, theta, ell + e)[0]
                - O.grad(cls, u, theta, ell - e)[0]) / (2 * h)
        assert gt[j] == pytest.approx(fd_t, rel=1e-5, abs=1e-4)
        assert ge[j] == pytest.approx(fd_e, rel=1e-5, abs=1e-4)
    # softmax invariance: the gradient is orthogonal to the ones vector
    assert abs(gt.sum()) < 1e-8 * np.abs(gt).sum()


def test_hvp_matches_gradient_differences():
    cls, u = tiny_problem(6, n_tx=25, n_cls=70)
    rng = np.random.default_rng(2)
    theta = rng.standard_normal(cls.n_tx)
    ell = np.log(cls.tx_len.astype(float))
    v = rng.standard_normal(cls.n_tx)
    hv = O.hvp(cls, u, theta, ell, v)
    h = 1e-5
    fd = (O.grad(cls, u, theta + h * v, ell)[1]
          - O.grad(cls, u, theta - h * v, ell)[1]) / (2 * h)
    np.testing.assert_allclose(hv, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    # symmetric: w.Hv == v.Hw
    w = rng.standard_normal(cls.n_tx)
    hw = O.hvp(cls, u, theta, ell, w)
    assert np.dot(w, hv) == pytest.approx(np.dot(v, hw), rel=1e-10)


# ---- monotonicity fallback (an option; neither reference has one) --------

FB_TABLES = (13, 23, 30)   # tables where hera's SQUAREM lowers ell


def fb_table(seed):
    return tiny_problem(seed, n_tx=60, n_cls=300, zero_frac=0.3)


def ll_path(cls, u, K, **kw):
    """ell after each of the first K iterations, and the iterates."""
    lam = np.full(cls.n_tx, 1.0 / cls.n_tx)
    out, lams, fbs = [], [], []
    for _ in range(K):
        r = O.estimate_count(cls, u, init=lam, min_round=1, max_round=1,
                             **kw)
        lam = r["lam"]
        lams.append(lam)
        fbs.append(int(r["fallback"][0]))
        out.append(O.loglik(cls, u, lam))
    return np.array(out), lams, fbs


@pytest.mark.parametrize("seed", FB_TABLES)
def test_fallback_keeps_loglik_monotone(seed):
    cls, u = fb_table(seed)
    ll0, _, _ = ll_path(cls, u, 30)
    # without the fallback hera's step lowers ell at some iteration
    assert np.any(np.diff(ll0) < -1e-9 * np.abs(ll0[1:]))
    ll1, lams, fbs = ll_path(cls, u, 30, fallback=1)
    assert sum(fbs) > 0
    assert np.all(np.diff(ll1) >= -1e-12 * np.abs(ll1[1:]))
    # a run of 30 iterations in one call takes the same decisions
    r = O.estimate_count(cls, u, min_round=30, max_round=30, fallback=1)
    np.testing.assert_array_equal(r["fallback"], fbs)
    np.testing.assert_array_equal(r["lam"], lams[-1])


@pytest.mark.parametrize("seed", FB_TABLES)
def test_fallback_iteration_ends_at_two_plain_maps(seed):
    # an iteration that falls back ends exactly at u2 = M(M(lam)), with
    # the stop test of the second map; one that does not is the plain
    # SQUAREM iteration
    cls, u = fb_table(seed)
    _, lams, fbs = ll_path(cls, u, 30, fallback=1)
    lam = np.full(cls.n_tx, 1.0 / cls.n_tx)
    seen = 0
    for k in range(30):
        u1, _ = O.em_map(cls, u, lam)
        u2, st2 = O.em_map(cls, u, u1)
        plain = O.estimate_count(cls, u, init=lam, min_round=1, max_round=1)
        fbr = O.estimate_count(cls, u, init=lam, min_round=1, max_round=1,
                               fallback=1)
        if fbs[k]:
            np.testing.assert_array_equal(lams[k], u2)
            assert fbr["stop"][0] == st2
            d, sc = O.dloglik(cls, u, plain["lam"], u2)
            assert d < -1e-10 * sc
            seen += 1
        else:
            np.testing.assert_array_equal(lams[k], plain["lam"])
            assert fbr["stop"][0] == plain["stop"][0]
        lam = lams[k]
    assert seen == sum(fbs) > 0


def test_fallback_off_is_the_reference_loop(cfg1):
    cls, u = cfg1
    a = O.estimate_count(cls, u)
    b = O.estimate_count(cls, u, fallback=0)
    np.testing.assert_array_equal(a["lam"], b["lam"])
    assert a["k"] == b["k"] and not b["fallback"].any()


def test_dloglik_is_the_loglik_difference():
    for seed in range(4):
        cls, u = tiny_problem(seed, zero_frac=0.2 if seed % 2 else 0.0)
        rng = np.random.default_rng(seed)
        lam = rng.uniform(1e-3, 1.0, cls.n_tx)
        mu = lam * rng.uniform(0.5, 2.0, cls.n_tx)
        d, sc = O.dloglik(cls, u, lam, mu)
        exact = O.loglik(cls, u, lam) - O.loglik(cls, u, mu)
        assert d == pytest.approx(exact, rel=1e-9, abs=1e-9 * sc)
        d2, sc2 = O.dloglik(cls, u, mu, lam)
        assert d2 == pytest.approx(-d, rel=1e-9, abs=1e-12 * sc)
        assert O.dloglik(cls, u, lam, lam)[0] == 0.0
        # a class whose members are all zero at one point only
        mu2 = mu.copy()
        m = cls.cls_ref[cls.cls_off[0]:cls.cls_off[1]]
        mu2[m] = 0.0
        u2 = u.copy()
        u2[m] = 0
        exact = O.loglik(cls, u2, lam) - O.loglik(cls, u2, mu2)
        assert O.dloglik(cls, u2, lam, mu2)[0] == pytest.approx(exact,
                                                               rel=1e-9)
