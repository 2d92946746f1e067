"""GPU adjoint, Hessian-vector products and standard errors.

The objective (include/hera_em.h, hem_grad) is the log-likelihood whose EM
map is hera's (src/EM.c:271-302).  Checks: the oracle's CPU gradient and
HVP (oracle_grad / oracle_hvp, themselves pinned by central differences in
test_oracle.py), FP64 central differences of the GPU's own log-likelihood
(h = 1e-6, bound 1e-5 relative), the stationarity of hera's EM fixed point,
and the CG standard errors against a dense pseudo-inverse of the Hessian
assembled from oracle HVPs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import hera_amd as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def sub():
    # the 2,000-transcript / 20,000-class validation subproblem
    cls, u = H.generate(2000, 20000, 1_000_000, seed=4)
    return cls, u


def point(cls, seed):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 1.0, cls.n_tx)
    ell = np.log(cls.tx_len.astype(float)) + rng.normal(0.0, 0.1, cls.n_tx)
    return theta, ell


def test_grad_matches_cpu_reference(sub):
    cls, u = sub
    theta, ell = point(cls, 1)
    p = H.Problem(cls)
    g = p.grad(u, theta, ell)
    p.close()
    L, gt, ge = O.grad(cls, u, theta, ell)
    say("loglik gpu %.12e cpu %.12e  N %.1f  %.3f ms" % (
        g["loglik"], L, g["n_frag"], g["ms"]))
    assert abs(g["loglik"] - L) <= 1e-12 * abs(L)
    assert np.max(np.abs(g["g_theta"] - gt)) <= 1e-10 * np.max(np.abs(gt))
    assert np.max(np.abs(g["g_log_efflen"] - ge)) <= 1e-12 * np.max(np.abs(ge))
    # softmax invariance: sum_t dL/dtheta_t = 0
    assert abs(g["g_theta"].sum()) <= 1e-9 * np.abs(g["g_theta"]).sum()


def test_grad_central_differences(sub):
    cls, u = sub
    theta, ell = point(cls, 2)
    p = H.Problem(cls)
    g = p.grad(u, theta, ell)
    h = 1e-6
    rng = np.random.default_rng(3)
    big = np.argsort(-np.abs(g["g_theta"]))[:20]
    idx = np.unique(np.concatenate([big, rng.choice(cls.n_tx, 20, False)]))
    fd_t, fd_e = [], []
    for j in idx:
        e = np.zeros(cls.n_tx)
        e[j] = h
        fd_t.append((p.grad(u, theta + e, ell)["loglik"]
                     - p.grad(u, theta - e, ell)["loglik"]) / (2 * h))
        fd_e.append((p.grad(u, theta, ell + e)["loglik"]
                     - p.grad(u, theta, ell - e)["loglik"]) / (2 * h))
    p.close()
    fd_t, fd_e = np.array(fd_t), np.array(fd_e)
    et = np.linalg.norm(fd_t - g["g_theta"][idx]) / np.linalg.norm(fd_t)
    ee = np.linalg.norm(fd_e - g["g_log_efflen"][idx]) / np.linalg.norm(fd_e)
    say("central differences h=1e-6: rel err theta %.2e, log_efflen %.2e"
        % (et, ee))
    assert et <= 1e-5 and ee <= 1e-5


def test_grad_fp32_forward(sub):
    cls, u = sub
    theta, ell = point(cls, 5)
    p = H.Problem(cls)
    g64 = p.grad(u, theta, ell)
    g32 = p.grad(u, theta, ell, precision=H.PREC_FP32)
    p.close()
    et = (np.linalg.norm(g32["g_theta"] - g64["g_theta"])
          / np.linalg.norm(g64["g_theta"]))
    ee = (np.linalg.norm(g32["g_log_efflen"] - g64["g_log_efflen"])
          / np.linalg.norm(g64["g_log_efflen"]))
    el = abs(g32["loglik"] - g64["loglik"]) / abs(g64["loglik"])
    say("FP32 forward vs FP64: theta %.2e, log_efflen %.2e, loglik %.2e"
        % (et, ee, el))
    assert et <= 1e-4 and ee <= 1e-5 and el <= 1e-6


def fixed_point(cls, u):
    """hera's EM run to a tight fixed point (every transcript with a unique
    read, so every lambda > 0); theta = log(lambda len), ell = log len."""
    p = H.Problem(cls)
    r = p.solve(u, stop_min_value=0.0, stop_rel_change=1e-13, min_iter=50,
                max_iter=20000, norm=H.NORM_TREE)
    p.close()
    lam = r["lam"]
    assert np.all(lam > 0)
    theta = np.log(lam * cls.tx_len)
    theta -= theta.mean()
    return theta, np.log(cls.tx_len.astype(float))


def test_grad_vanishes_at_em_fixed_point(sub):
    cls, u = sub
    u1 = u + 1
    theta, ell = fixed_point(cls, u1)
    p = H.Problem(cls)
    g = p.grad(u1, theta, ell)
    gt = O.grad(cls, u1, theta, ell)
    p.close()
    E = -g["g_log_efflen"]
    rel = np.max(np.abs(g["g_theta"]) / (E + 1.0))
    say("max |dL/dtheta| / (E + 1) at the fixed point: %.2e" % rel)
    assert rel <= 1e-6
    np.testing.assert_allclose(g["g_log_efflen"], gt[2], rtol=1e-12)


def test_hvp_matches_cpu_and_gradient_differences(sub):
    cls, u = sub
    theta, ell = point(cls, 6)
    rng = np.random.default_rng(7)
    v = rng.normal(size=cls.n_tx)
    p = H.Problem(cls)
    hv = p.hvp(u, theta, ell, v)
    h = 1e-6
    fd = (p.grad(u, theta + h * v, ell)["g_theta"]
          - p.grad(u, theta - h * v, ell)["g_theta"]) / (2 * h)
    p.close()
    ref = O.hvp(cls, u, theta, ell, v)
    e_cpu = np.linalg.norm(hv - ref) / np.linalg.norm(ref)
    e_fd = np.linalg.norm(hv - fd) / np.linalg.norm(fd)
    say("HVP vs CPU %.2e, vs central differences of the gradient %.2e"
        % (e_cpu, e_fd))
    assert e_cpu <= 1e-11
    assert e_fd <= 1e-5
    # the ones vector is in the null space (softmax shift invariance)
    p = H.Problem(cls)
    h1 = p.hvp(u, theta, ell, np.ones(cls.n_tx))
    p.close()
    assert np.max(np.abs(h1)) <= 1e-8 * np.max(np.abs(hv))


def test_grad_and_hvp_on_untiled_table():
    """A component over the tile limits (a chain of 100 transcripts) sends
    the adjoint and the HVP to the global CSR/CSC passes; both still equal
    the CPU reference."""
    rng = np.random.default_rng(11)
    T = 300
    members = [[i, i + 1] for i in range(99)]
    members += [sorted(rng.choice(np.arange(100, T), rng.integers(1, 5),
                                  replace=False).tolist())
                for _ in range(600)]
    counts = rng.integers(1, 50, len(members)).astype(float).tolist()
    cls = H.Classes.from_lists(T, members, counts,
                               rng.integers(200, 3000, T).tolist())
    u = rng.integers(0, 20, T).astype(np.uint32)
    theta, ell = point(cls, 12)
    v = rng.normal(size=T)
    p = H.Problem(cls)
    try:
        assert p.tiles()["n_tiles"] == 0
        g = p.grad(u, theta, ell)
        hv = p.hvp(u, theta, ell, v)
    finally:
        p.close()
    L, gt, ge = O.grad(cls, u, theta, ell)
    assert abs(g["loglik"] - L) <= 1e-12 * abs(L)
    assert np.max(np.abs(g["g_theta"] - gt)) <= 1e-10 * np.max(np.abs(gt))
    ref = O.hvp(cls, u, theta, ell, v)
    assert np.linalg.norm(hv - ref) <= 1e-11 * np.linalg.norm(ref)


def dense_fisher(cls, u, theta, ell):
    T = cls.n_tx
    Hm = np.empty((T, T))
    for j in range(T):
        e = np.zeros(T)
        e[j] = 1.0
        Hm[:, j] = O.hvp(cls, u, theta, ell, e)
    return -(Hm + Hm.T) / 2


def test_stderr_against_dense_pseudo_inverse(sub):
    cls, u = sub
    u1 = u + 1
    theta, ell = fixed_point(cls, u1)
    K = 64
    p = H.Problem(cls)
    long_run = p.stderr(u1, theta, ell, top_k=K, cg_steps=400)
    short = p.stderr(u1, theta, ell, top_k=K, cg_steps=50)
    again = p.stderr(u1, theta, ell, top_k=K, cg_steps=50)
    p.close()
    F = dense_fisher(cls, u1, theta, ell)
    pi = np.exp(theta - theta.max())
    pi /= pi.sum()
    order = sorted(range(cls.n_tx), key=lambda t: (-pi[t], t))[:K]
    np.testing.assert_array_equal(long_run["idx"], order)
    Fp = np.linalg.pinv(F, rcond=1e-13, hermitian=True)
    se = []
    for j in order:
        b = -pi[j] * pi
        b[j] += pi[j]
        se.append(np.sqrt(max(b @ Fp @ b, 0.0)))
    se = np.array(se)
    e_long = np.max(np.abs(long_run["se"] - se) / se)
    e_short = np.max(np.abs(short["se"] - se) / se)
    say("SE vs dense pseudo-inverse: 400 CG steps %.2e (resid %.1e), "
        "50 steps %.2e (resid %.1e), %.1f ms"
        % (e_long, long_run["stats"]["resid_max"], e_short,
           short["stats"]["resid_max"], short["stats"]["ms_cg"]))
    assert e_long <= 1e-6
    assert e_short <= 1e-3
    np.testing.assert_array_equal(short["se"], again["se"])


def test_autograd_function_gradcheck():
    """torch.autograd.Function around hem_grad (test harness only), checked
    by torch.autograd.gradcheck in FP64 in its own process."""
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.dirname(HERE) + os.pathsep + env.get(
        "PYTHONPATH", "")
    r = subprocess.run([sys.executable, "-u",
                        os.path.join(HERE, "autograd_harness.py")],
                       env=env, capture_output=True, text=True, timeout=600)
    say(r.stdout[-2000:])
    say(r.stderr[-2000:])
    assert r.returncode == 0
    assert "gradcheck ok" in r.stdout
