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


def hera_em_reads(seed, n_tx=30, n_pat=200, max_aln=4):
    """A hera-EM simple-model read set: n_pat distinct read patterns, each a
    list of (transcript, c_prob) alignments, a read length and a count of
    identical reads."""
    rng = np.random.default_rng(seed)
    pats = []
    for _ in range(n_pat):
        k = int(rng.integers(1, max_aln + 1))
        ts = rng.choice(n_tx, size=k, replace=False)
        cp = rng.uniform(1e-3, 1.0, size=k)
        pats.append(([(int(t), float(c)) for t, c in zip(ts, cp)],
                     int(rng.integers(3, 9)), int(rng.integers(1, 60))))
    return pats


def hera_em_table(n_tx, pats):
    """The class table whose weighted map is hera-EM's simple-model E-step:
    one class per read pattern (count = identical reads), its alignments as
    members weighted by c_prob, and the noise as transcript n_tx, a member
    of every class with weight 4^-read_length (hera-EM/smodel.h:15-26,
    utility.c:10-12).  Lengths 1, no unique reads."""
    from hera_amd import Classes
    members, weights, counts = [], [], []
    for aln, rl, n in pats:
        members.append([t for t, _ in aln] + [n_tx])
        weights += [c for _, c in aln] + [0.25 ** rl]
        counts.append(float(n))
    c = Classes.from_lists(n_tx + 1, members, counts, [1] * (n_tx + 1))
    c.member_weight = np.array(weights)
    return c, np.zeros(n_tx + 1, np.uint32)
