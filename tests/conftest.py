import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950); runs the HIP path")


def tiny_problem(seed, n_tx=40, n_cls=120, max_size=6, dup=False,
                 zero_frac=0.0):
    """Small random class table for CPU-sized parity checks.

    Members are drawn anywhere in [0, n_tx); with dup=True some classes
    repeat a member (hera's concatenated paired candidates).
    """
    from hera_amd import Classes
    rng = np.random.default_rng(seed)
    members, counts = [], []
    for _ in range(n_cls):
        s = int(rng.integers(2, max_size + 1))
        m = rng.choice(n_tx, size=s, replace=False)
        if dup and rng.random() < 0.3:
            m = np.concatenate([m, m[:1]])
        members.append(np.sort(m) if not dup else m)
        counts.append(float(rng.integers(0 if zero_frac else 1, 500)))
    lens = rng.integers(200, 5000, size=n_tx)
    unique = rng.integers(0, 300, size=n_tx)
    if zero_frac:
        unique[rng.random(n_tx) < zero_frac] = 0
    return Classes.from_lists(n_tx, members, counts, lens), unique.astype(
        np.uint32)


@pytest.fixture(scope="session")
def cfg1():
    """Config 1 of the benchmark generator: 20k transcripts, 200k classes
    drawn (singletons folded into unique counts), 10M fragments, seed 0."""
    import hera_amd as H
    return H.generate(20000, 200000, 10_000_000, seed=0)
