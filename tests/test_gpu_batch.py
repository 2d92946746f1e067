"""GPU parity of batched solves and the bootstrap against the FP64 oracle.

Every replicate of a batch is an independent hera EM on its own counts
(src/EM.c:315-387); in the default mode (FP64, exact norms) each must equal
the oracle run on the same counts bit for bit, iteration count included.
The bootstrap's resampled counts must equal the oracle's restatement of the
resampler bit for bit.
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


def with_counts(cls, counts):
    return H.Classes(cls.n_tx, cls.cls_off, cls.cls_ref,
                     np.asarray(counts, dtype=np.float64), cls.tx_len)


def check_replicates(cls, U, CC, res, bitwise=True, tol=0.0, **kw):
    for r in range(U.shape[0]):
        ora = O.estimate_count(with_counts(cls, CC[r]), U[r], **kw)
        assert res["iters"][r] == ora["k"] or not bitwise, (r, res["iters"][r],
                                                           ora["k"])
        if bitwise:
            np.testing.assert_array_equal(res["est"][r], ora["est"])
            np.testing.assert_array_equal(res["tpm"][r], ora["tpm"])
        else:
            m = ora["tpm"] / 1e6 > 1e-7
            e = np.abs(res["tpm"][r][m] - ora["tpm"][m]) / ora["tpm"][m]
            say("replicate %d k=%d/%d max err %.3e" % (
                r, res["iters"][r], ora["k"], e.max()))
            assert e.max() <= tol


def perturbed(cls, u, R, seed):
    rng = np.random.default_rng(seed)
    U = np.array([rng.poisson(u) for _ in range(R)], dtype=np.uint32)
    CC = np.array([rng.poisson(cls.cls_count) for _ in range(R)],
                  dtype=np.float64)
    return U, CC


@pytest.mark.parametrize("R", [1, 3, 17])
def test_batch_tiny_bitwise(R):
    cls, u = tiny_problem(R, n_tx=50, n_cls=200, dup=True)
    U, CC = perturbed(cls, u, R, 100 + R)
    p = H.Problem(cls)
    assert p.tiles()["n_tiles"] >= 1
    res = p.solve_batch(U, CC)
    say("R=%d iters %s chunk %d" % (R, res["iters"].tolist(),
                                    res["stats"]["chunk"]))
    check_replicates(cls, U, CC, res)
    # R = 1 batch equals the plain solve
    if R == 1:
        g = p.solve(U[0]) if np.array_equal(CC[0], cls.cls_count) else None
        if g is not None:
            np.testing.assert_array_equal(res["est"][0], g["est"])
    p.close()


def test_batch_shared_counts_equals_plain(cfg1):
    cls, u = cfg1
    p = H.Problem(cls)
    U = np.stack([u, u])
    res = p.solve_batch(U)   # table's own class counts for both replicates
    g = p.solve(u)
    p.close()
    np.testing.assert_array_equal(res["est"][0], g["est"])
    np.testing.assert_array_equal(res["est"][1], g["est"])
    np.testing.assert_array_equal(res["tpm"][0], g["tpm"])
    assert res["iters"][0] == g["stats"]["iters"]


@pytest.mark.parametrize("kw", [dict(), dict(step_form=1),
                                dict(clamp_form=1), dict(stop_rel=1e-8)])
def test_batch_cfg1_bitwise(cfg1, kw):
    cls, u = cfg1
    R = 6
    U, CC = perturbed(cls, u, R, 7)
    gmap = dict(step_form="step", clamp_form="clamp",
                stop_rel="stop_rel_change")
    p = H.Problem(cls)
    res = p.solve_batch(U, CC, **{gmap[k]: v for k, v in kw.items()})
    p.close()
    say("cfg1 %s iters %s ms %.1f" % (kw, res["iters"].tolist(),
                                      res["stats"]["ms_solve"]))
    check_replicates(cls, U, CC, res, **kw)


@pytest.mark.parametrize("step", [H.STEP_SD, H.STEP_L2])
def test_batch_exact_sums_equal_single_lane_chains(cfg1, step):
    """Speculated exact sums and one-lane chains: the same batch, bit for
    bit (replicates finish at different iterations, so the active list
    shrinks under both)."""
    cls, u = cfg1
    U, CC = perturbed(cls, u, 5, 21)
    p = H.Problem(cls)
    a = p.solve_batch(U, CC, norm=H.NORM_EXACT, step=step)
    b = p.solve_batch(U, CC, norm=H.NORM_SEQ, step=step)
    p.close()
    say("step %d iters %s reruns %d" % (step, a["iters"].tolist(),
                                        a["stats"]["sum_reruns"]))
    np.testing.assert_array_equal(a["iters"], b["iters"])
    np.testing.assert_array_equal(a["est"], b["est"])
    np.testing.assert_array_equal(a["tpm"], b["tpm"])


def test_batch_tree_norms(cfg1):
    # parallel (Chan-combined) norms put alpha within an ulp of the
    # reference's; the trajectory amplifies that to ~1e-6..1e-5 on weakly
    # identifiable abundances, so this mode is held to 1e-4, not 1e-6
    cls, u = cfg1
    U, CC = perturbed(cls, u, 4, 9)
    p = H.Problem(cls)
    res = p.solve_batch(U, CC, norm=H.NORM_TREE)
    p.close()
    check_replicates(cls, U, CC, res, bitwise=False, tol=1e-4)


def test_bootstrap_counts_bitwise(cfg1):
    cls, u = cfg1
    seeds = np.array([11, 12, 2 ** 63 + 5], dtype=np.uint64)
    p = H.Problem(cls)
    uo, co = p.bootstrap_counts(u, seeds)
    p.close()
    w = np.concatenate([u, cls.cls_count.astype(np.uint32)])
    N = int(w.sum(dtype=np.uint64))
    for r, s in enumerate(seeds):
        ref = O.resample(w, int(s))
        np.testing.assert_array_equal(uo[r], ref[:cls.n_tx])
        np.testing.assert_array_equal(co[r], ref[cls.n_tx:])
        assert int(uo[r].sum(dtype=np.uint64)) + int(co[r].sum()) == N


def test_bootstrap_end_to_end_bitwise(cfg1):
    cls, u = cfg1
    seeds = np.arange(5, dtype=np.uint64) * 7919 + 3
    p = H.Problem(cls)
    res = p.bootstrap(u, seeds)
    uo, co = p.bootstrap_counts(u, seeds)
    again = p.bootstrap(u, seeds)
    p.close()
    say("bootstrap iters %s resample %.2f ms solve %.2f ms" % (
        res["iters"].tolist(), res["stats"]["ms_resample"],
        res["stats"]["ms_solve"]))
    np.testing.assert_array_equal(res["est"], again["est"])
    check_replicates(cls, uo, co, res)


def test_bootstrap_after_plain_solves_on_same_handle(cfg1):
    """Plain solves (tile kernels, exact sums, graphs) and then bootstraps
    on one handle give what a fresh handle gives, bit for bit."""
    cls, u = cfg1
    seeds = np.arange(3, dtype=np.uint64) + 5
    p = H.Problem(cls)
    for norm in (H.NORM_EXACT, H.NORM_TREE, H.NORM_SEQ):
        p.solve(u, norm=norm, stop_rel_change=1e-8)
    p.solve(u, precision=H.PREC_FP32, norm=H.NORM_TREE, max_iter=60)
    a = p.bootstrap(u, seeds[:2])
    a = p.bootstrap(u, seeds)
    b = p.bootstrap(u, seeds, norm=H.NORM_TREE)
    g = p.solve(u)
    p.close()
    q = H.Problem(cls)
    a2 = q.bootstrap(u, seeds)
    b2 = q.bootstrap(u, seeds, norm=H.NORM_TREE)
    g2 = q.solve(u)
    q.close()
    for x, y in ((a, a2), (b, b2)):
        np.testing.assert_array_equal(x["iters"], y["iters"])
        np.testing.assert_array_equal(x["est"], y["est"])
        np.testing.assert_array_equal(x["tpm"], y["tpm"])
    np.testing.assert_array_equal(g["lam"], g2["lam"])


@pytest.mark.parametrize("case", ["cfg1", "zero_component"])
def test_batch_fp32_equals_plain_fp32(cfg1, case):
    """The FP32 bulk form batched: each replicate is hem_solve's FP32 solve
    of its own counts, bit for bit, with exact norms (the same arithmetic
    on small and large tiles), iteration count included."""
    if case == "cfg1":
        cls, u = cfg1
        U, CC = perturbed(cls, u, 4, 31)
    else:
        cls, u = zero_component_problem(5)
        U, CC = perturbed(cls, u, 3, 32)
        U[:, -4:] = 0
        CC[:, -3:] = 0
    kw = dict(precision=H.PREC_FP32, norm=H.NORM_EXACT,
              stop_rel_change=1e-6, max_iter=300)
    p = H.Problem(cls)
    res = p.solve_batch(U, CC, **kw)
    tree = p.solve_batch(U, CC, precision=H.PREC_FP32, norm=H.NORM_TREE,
                         stop_rel_change=1e-6, max_iter=300)
    p.close()
    say("%s fp32 iters %s (tree %s)" % (case, res["iters"].tolist(),
                                        tree["iters"].tolist()))
    for r in range(U.shape[0]):
        cr = with_counts(cls, CC[r])
        q = H.Problem(cr)
        g = q.solve(U[r], **kw)
        q.close()
        assert res["iters"][r] == g["stats"]["iters"]
        np.testing.assert_array_equal(res["est"][r], g["est"])
        np.testing.assert_array_equal(res["tpm"][r], g["tpm"])
        # the FP32 form against the FP64 oracle on the same counts, held
        # to the log-likelihood bar of that mode
        ora = O.estimate_count(cr, U[r], stop_rel=1e-6, max_round=300)
        l32 = O.loglik(cr, U[r], g["lam"])
        l64 = O.loglik(cr, U[r], ora["lam"])
        say("replicate %d: k %d (oracle %d), loglik rel diff %.2e" % (
            r, g["stats"]["iters"], ora["k"], abs(l32 - l64) / abs(l64)))
        assert abs(l32 - l64) <= 1e-9 * abs(l64)
        assert np.all(np.isfinite(tree["est"][r]))


@pytest.mark.parametrize("k", [50, 80, 100])
def test_batch_tiny_densities_bitwise(k):
    """Pairs {A, B} where only A has unique reads: under hera-EM's clamp
    B's density decays through (0, 2^-299) to denormals and 0 within 100
    iterations (checked on the oracle below: at 50, 80 and 100 iterations
    some replicate ends with densities in (0, 2^-299), at 100 with
    denormals), so the batched kernel's tiles leave its fast quotient for
    the division; every replicate stays the oracle's, bit for bit, next to
    tiles of ordinary densities."""
    P = 40
    members = [[2 * i, 2 * i + 1] for i in range(P)]
    counts = [float(50 + 3 * i) for i in range(P)]
    big, ub = tiny_problem(9, n_tx=60, n_cls=240, dup=True)
    T0 = big.n_tx
    members = [big.cls_ref[big.cls_off[i]:big.cls_off[i + 1]].tolist()
               for i in range(big.n_cls)] + [[T0 + a for a in m]
                                             for m in members]
    counts = big.cls_count.tolist() + counts
    lens = big.tx_len.tolist() + [1000 + 10 * i for i in range(2 * P)]
    cls = H.Classes.from_lists(T0 + 2 * P, members, counts, lens)
    u = np.concatenate([ub, np.zeros(2 * P, np.uint32)])
    u[T0::2] = 1000 + np.arange(P)
    U, CC = perturbed(cls, u, 3, 40 + k)
    U[:, T0 + 1::2] = 0
    kw = dict(step_form=1, clamp_form=1, stop_rel=0.0, min_round=k,
              max_round=k)
    tiny = False
    for r in range(U.shape[0]):
        b = O.estimate_count(with_counts(cls, CC[r]), U[r],
                             **kw)["lam"][T0 + 1::2]
        tiny |= bool(np.any((b > 0) & (b < 2.0 ** -299)))
    assert tiny
    p = H.Problem(cls)
    try:
        res = p.solve_batch(U, CC, step=1, clamp=1, stop_rel_change=0.0,
                            min_iter=k, max_iter=k)
    finally:
        p.close()
    check_replicates(cls, U, CC, res, **kw)


def test_batch_rejects_untileable_and_unknown_precision():
    members = [[i, i + 1] for i in range(99)]
    big = H.Classes.from_lists(100, members, [3.0] * 99, [500] * 100)
    p = H.Problem(big)
    assert p.tiles()["n_tiles"] == 0
    with pytest.raises(H.HemError) as ei:
        p.solve_batch(np.ones((2, 100), np.uint32))
    assert ei.value.code == _lib.HEM_ERR_RANGE
    p.close()
    cls, u = tiny_problem(1)
    p = H.Problem(cls)
    with pytest.raises(H.HemError) as ei:
        p.solve_batch(np.stack([u, u]), precision=7)
    assert ei.value.code == _lib.HEM_ERR_INVALID
    p.close()


def zero_component_problem(seed):
    """A generated table plus a component of four transcripts whose unique
    and class counts are all 0: after the first map its densities are exact
    zeros and its classes have D_c == 0, which the reference skips
    (src/EM.c:279-280), beside components that carry reads."""
    cls, u = tiny_problem(seed, n_tx=60, n_cls=240, dup=True)
    T = cls.n_tx
    members = [cls.cls_ref[cls.cls_off[i]:cls.cls_off[i + 1]].tolist()
               for i in range(cls.n_cls)]
    members += [[T, T + 1], [T + 1, T + 2, T + 3], [T + 3, T]]
    counts = np.concatenate([cls.cls_count, [0.0, 0.0, 0.0]])
    lens = np.concatenate([cls.tx_len, [300, 400, 500, 600]])
    big = H.Classes.from_lists(T + 4, members, counts.tolist(), lens.tolist())
    return big, np.concatenate([u, np.zeros(4, np.uint32)])


def test_zero_density_classes_bitwise():
    cls, u = zero_component_problem(4)
    p = H.Problem(cls)
    assert p.tiles()["n_tiles"] > 0
    g = p.solve(u)
    ora = O.estimate_count(cls, u)
    assert g["stats"]["iters"] == ora["k"]
    np.testing.assert_array_equal(g["lam"], ora["lam"])
    assert (g["lam"][-4:] == 0).all()
    U, CC = perturbed(cls, u, 5, 77)
    CC[:, -3:] = 0
    U[:, -4:] = 0
    res = p.solve_batch(U, CC)
    p.close()
    check_replicates(cls, U, CC, res)


def limit_component_problem(seed):
    """A generated table plus one component at the tile limits: 64
    transcripts and 320 classes (a chain through all 64 transcripts, then
    random classes inside the component), so one tile of each tiling holds
    it alone and sets the kernels' LDS allocation."""
    cls, u = H.generate(400, 4000, 200_000, seed=seed)
    rng = np.random.default_rng(seed)
    T = cls.n_tx
    members = [cls.cls_ref[cls.cls_off[i]:cls.cls_off[i + 1]].tolist()
               for i in range(cls.n_cls)]
    comp = [[T + i, T + i + 1] for i in range(63)]
    while len(comp) < 320:
        k = int(rng.integers(2, 9))
        comp.append((T + rng.choice(64, size=k, replace=False)).tolist())
    members += comp
    counts = np.concatenate([cls.cls_count,
                             rng.integers(1, 400, size=320).astype(float)])
    lens = np.concatenate([cls.tx_len, rng.integers(200, 5000, size=64)])
    u2 = np.concatenate([u, rng.integers(0, 300, size=64).astype(np.uint32)])
    return H.Classes.from_lists(T + 64, members, counts.tolist(),
                                lens.tolist()), u2


def test_tile_limit_component_bitwise():
    cls, u = limit_component_problem(6)
    t = cls.tiling()
    assert t["n_tiles"] > 0 and t["max_tx"] == 64 and t["max_cls"] == 320
    p = H.Problem(cls)
    g = p.solve(u)
    ora = O.estimate_count(cls, u)
    assert g["stats"]["iters"] == ora["k"]
    np.testing.assert_array_equal(g["lam"], ora["lam"])
    U, CC = perturbed(cls, u, 3, 91)
    res = p.solve_batch(U, CC)
    p.close()
    say("limit tile: chunk %d" % res["stats"]["chunk"])
    check_replicates(cls, U, CC, res)


def test_profile_batch_follows_the_solve(cfg1):
    """hem_profile_batch runs the solves of hem_solve_batch launch by
    launch: as many iterations as the slowest replicate's k, as many
    replicate-iterations as all k together, one extrapolated-map launch
    per iteration and the first maps once."""
    cls, u = cfg1
    U, CC = perturbed(cls, u, 6, 77)
    p = H.Problem(cls)
    try:
        res = p.solve_batch(U, CC, stop_rel_change=1e-6)
        pr = p.profile_batch(U, CC, stop_rel_change=1e-6)
        pt = p.profile_batch(U, CC, stop_rel_change=1e-6, norm=H.NORM_TREE)
        rt = p.solve_batch(U, CC, stop_rel_change=1e-6, norm=H.NORM_TREE)
    finally:
        p.close()
    it = res["iters"]
    say("iters %s; profile %d iterations, %d replicate-iterations" % (
        it.tolist(), pr["iters"], pr["rep_iters"]))
    for prof, its in ((pr, it), (pt, rt["iters"])):
        assert prof["iters"] == its.max()
        assert prof["rep_iters"] == its.sum()
        assert prof["tile_map12"]["launches"] == 1
        assert prof["tile_ex3"]["launches"] == prof["iters"]
        assert prof["iter_end"]["launches"] == prof["iters"]
        assert prof["tile_ex3"]["ms_total"] > 0
        # 3 maps per launch: 3 (8Z + 4C + 8T) per launch plus
        # 3 (12C + 16T) per replicate-iteration
        T, C, Z = cls.n_tx, cls.n_cls, cls.nnz
        exp = (3 * (8 * Z + 4 * C + 8 * T) * prof["iters"]
               + 3 * (12 * C + 16 * T) * prof["rep_iters"])
        assert prof["tile_ex3"]["bytes_total"] == pytest.approx(exp,
                                                                rel=1e-12)
    assert pr["norm_pass2"]["launches"] == pr["iters"]
