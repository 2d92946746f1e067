"""hera's abundance tables (hem_write_abundance, hem_write_gene_abundance,
hem_write_result) against the oracle's restatement of write_result
(src/EM.c:152-205), byte for byte, and against what the format itself
fixes: the header, six tab-separated fields, values to %f's six decimals,
gene sums accumulated in float, run boundaries and the reference's file
names.  Host code only: no device is used.
"""
import os

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from oracle import oracle as O


def table(n, seed, name_len=40):
    rng = np.random.default_rng(seed)
    names = ["ENST%011d.%d|%s" % (rng.integers(1e10), rng.integers(1, 9),
                                  "x" * int(rng.integers(0, 12)))
             for _ in range(n)]
    unique = rng.integers(0, 5_000_000, n).astype(np.uint32)
    length = rng.integers(200, 20000, n).astype(np.uint32)
    eff = length - rng.uniform(0, 180, n)
    est = rng.exponential(50.0, n)
    est[rng.random(n) < 0.2] = 0.0
    est[:3] = [1e12 / 3, 1e-12, 123456.7890125]
    tpm = est / est.sum() * 1e6
    return names, unique, length, eff, est, tpm, name_len


def genes(n, seed):
    rng = np.random.default_rng(seed)
    gid = np.repeat(np.arange(n), rng.integers(1, 6, n))[:n]
    # ids need not be sorted; runs of equal ids are what is grouped
    perm = rng.permutation(gid.max() + 1)
    gid = perm[gid].astype(np.uint32)
    gnames = ["ENSG%011d.%d|G%d" % (i * 7919, i % 9, i)
              for i in range(gid.max() + 1)]
    return gid, gnames


def read(p):
    with open(p, "rb") as f:
        return f.read()


def test_abundance_tsv_equals_oracle(tmp_path):
    names, u, ln, eff, est, tpm, w = table(3000, 1)
    blk, w = H.name_blocks(names, w)
    a, b = tmp_path / "a.tsv", tmp_path / "b.tsv"
    H.write_abundance(a, names, u, ln, eff, est, tpm, name_len=w)
    O.write_abundance(b, blk, w, u, ln, eff, est, tpm)
    assert read(a) == read(b)


def test_abundance_tsv_fields(tmp_path):
    names, u, ln, eff, est, tpm, w = table(500, 2)
    p = tmp_path / "a.tsv"
    H.write_abundance(p, names, u, ln, eff, est, tpm, name_len=w)
    lines = open(p).read().split("\n")
    assert lines[0] == ("#target_id\tunique_map\tlength\teff_length\t"
                        "est_counts\ttpm")
    assert lines[-1] == "" and len(lines) == 502
    for i, line in enumerate(lines[1:-1]):
        f = line.split("\t")
        assert len(f) == 6
        assert f[0] == names[i]
        assert int(f[1]) == u[i] and int(f[2]) == ln[i]
        for k, v in ((3, eff[i]), (4, est[i]), (5, tpm[i])):
            # %f: six decimals, rounded
            assert len(f[k].split(".")[1]) == 6
            assert abs(float(f[k]) - v) <= 5e-7 + 1e-15 * abs(v)


def test_gene_tsv_equals_oracle(tmp_path):
    n = 2000
    names, u, ln, eff, est, tpm, w = table(n, 3)
    gid, gnames = genes(n, 4)
    blk, w = H.name_blocks(names, w)
    gblk, gw = H.name_blocks(gnames)
    a, b = tmp_path / "g1.tsv", tmp_path / "g2.tsv"
    H.write_gene_abundance(a, names, gid, gnames, est, tpm, name_len=w,
                           gene_len=gw)
    O.write_gene_abundance(b, blk, w, gid, gblk, gw, est, tpm)
    assert read(a) == read(b)


def test_gene_runs_float_sums_and_names(tmp_path):
    n = 700
    names, u, ln, eff, est, tpm, w = table(n, 5)
    gid, gnames = genes(n, 6)
    p = tmp_path / "g.tsv"
    H.write_gene_abundance(p, names, gid, gnames, est, tpm, name_len=w)
    lines = open(p).read().split("\n")[1:-1]
    # runs of consecutive equal ids, float accumulation (src/EM.c:190-199)
    rows, i = [], 0
    while i < n:
        s0, s1 = np.float32(est[i]), np.float32(tpm[i])
        while i + 1 < n and gid[i + 1] == gid[i]:
            i += 1
            s0 = np.float32(np.float64(s0) + est[i])
            s1 = np.float32(np.float64(s1) + tpm[i])
        rows.append((names[i][16:], gnames[gid[i]][16:], s0, s1))
        i += 1
    assert len(lines) == len(rows)
    for line, (nm, gn, s0, s1) in zip(lines, rows):
        f = line.split("\t")
        assert f[0] == nm and f[1] == gn
        assert f[2] == "%f" % float(s0) and f[3] == "%f" % float(s1)
    # double accumulation would differ somewhere on this table
    dbl = []
    i = 0
    while i < n:
        s = est[i]
        while i + 1 < n and gid[i + 1] == gid[i]:
            i += 1
            s += est[i]
        dbl.append("%f" % s)
        i += 1
    assert [l.split("\t")[2] for l in lines] != dbl


def test_write_result_names(tmp_path):
    names, u, ln, eff, est, tpm, w = table(50, 7)
    gid, gnames = genes(50, 8)
    d = tmp_path / "out"
    d.mkdir()
    H.write_result(str(d), "s1", names, u, ln, eff, est, tpm, gene=gid,
                   gene_names=gnames, name_len=w)
    assert (d / "s1_abundance.tsv").exists()
    assert (d / "s1_abundance.gene.tsv").exists()
    H.write_result(str(d) + "/", "", names, u, ln, eff, est, tpm,
                   name_len=w)
    assert (d / "abundance.tsv").exists()
    assert not (d / "abundance.gene.tsv").exists()
    H.write_abundance(tmp_path / "x.tsv", names, u, ln, eff, est, tpm,
                      name_len=w)
    assert read(d / "abundance.tsv") == read(tmp_path / "x.tsv")


def test_writer_errors(tmp_path):
    names, u, ln, eff, est, tpm, w = table(10, 9)
    gid = np.zeros(10, np.uint32)
    with pytest.raises(H.HemError) as ei:
        H.write_gene_abundance(tmp_path / "g.tsv", names, gid, ["g"],
                               est, tpm, name_len=w)   # gene block <= 16
    assert ei.value.code == _lib.HEM_ERR_INVALID
    with pytest.raises(H.HemError) as ei:
        H.write_gene_abundance(tmp_path / "g.tsv", names, gid + 3,
                               ["G" * 20], est, tpm, name_len=w)
    assert ei.value.code == _lib.HEM_ERR_INVALID
    with pytest.raises(H.HemError) as ei:
        H.write_abundance(tmp_path / "no" / "such" / "a.tsv", names, u, ln,
                          eff, est, tpm, name_len=w)
    assert ei.value.code == _lib.HEM_ERR_IO
    H.write_abundance(tmp_path / "e.tsv", [], [], [], [], [], [])
    assert read(tmp_path / "e.tsv").count(b"\n") == 1
