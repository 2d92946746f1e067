"""torch.autograd.Function around the C-ABI adjoint (test harness only).

Run as a script by tests/test_gpu_grad.py: builds a small problem, wraps
hem_grad's log-likelihood and gradient as an autograd Function of (theta,
log_efflen) and checks it with torch.autograd.gradcheck (FP64 central
differences against the analytic gradient).
"""
import sys

import numpy as np
import torch

import hera_amd as H


class LogLik(torch.autograd.Function):
    """L(theta, log_efflen) of include/hera_em.h; backward from hem_grad."""

    @staticmethod
    def forward(ctx, theta, ell, prob, unique):
        g = prob.grad(unique, theta.detach().numpy(), ell.detach().numpy())
        ctx.save_for_backward(torch.from_numpy(g["g_theta"]),
                              torch.from_numpy(g["g_log_efflen"]))
        return theta.new_tensor(g["loglik"])

    @staticmethod
    def backward(ctx, gout):
        gt, ge = ctx.saved_tensors
        return gout * gt, gout * ge, None, None


def main():
    cls, u = H.generate(300, 3000, 200_000, seed=9)
    prob = H.Problem(cls)
    rng = np.random.default_rng(1)
    theta = torch.tensor(rng.normal(0, 1, cls.n_tx), dtype=torch.float64,
                         requires_grad=True)
    ell = torch.tensor(np.log(cls.tx_len.astype(float))
                       + rng.normal(0, 0.1, cls.n_tx),
                       dtype=torch.float64, requires_grad=True)
    # backward through the Function equals hem_grad
    L = LogLik.apply(theta, ell, prob, u)
    L.backward()
    g = prob.grad(u, theta.detach().numpy(), ell.detach().numpy())
    assert np.array_equal(theta.grad.numpy(), g["g_theta"])
    # |L| ~ 1e6: central differences at eps 1e-6 carry ~1e-4 of rounding
    ok = torch.autograd.gradcheck(
        lambda a, b: LogLik.apply(a, b, prob, u), (theta, ell), eps=1e-6,
        atol=2e-3, rtol=1e-5, nondet_tol=0.0, fast_mode=False)
    prob.close()
    print("gradcheck ok" if ok else "gradcheck failed", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
