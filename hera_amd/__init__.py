"""hera_amd: MI355X-native SQUAREM EM abundance estimation for hera.

The product is the C library ``libheraem.so`` (include/hera_em.h and the
drop-in ``estimate_count`` of include/hera_dropin.h).  This package is a thin
ctypes view of it for tests and benchmarks.  Nothing here computes on the
CPU: every call goes to the HIP path and raises if the library is missing.
"""
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import HemError, as_f64, as_u32, check, f64p, u32p

__all__ = ["Classes", "Problem", "generate", "opts", "HemError",
           "PREC_FP32", "PREC_FP64"]

PREC_FP32 = 0
PREC_FP64 = 1
STEP_SD = 0
STEP_L2 = 1
CLAMP_ZERO = 0
CLAMP_DIV10 = 1
NORM_EXACT = 0
NORM_TREE = 1
NORM_SEQ = 2
NORM_EXACT_WIN = 3


@dataclass
class Classes:
    """hera's class table in CSR form (Gclass + Ref_inf->len).

    Class c has members cls_ref[cls_off[c]:cls_off[c+1]] and cls_count[c]
    fragments; tx_len[t] is the u32 transcript length.  member_weight
    (optional, one float per member, >= 0) weighs each member's share of its
    class as hera-EM's alignment probabilities do (hera-EM/smodel.h:15-20);
    None is hera's model.
    """
    n_tx: int
    cls_off: np.ndarray
    cls_ref: np.ndarray
    cls_count: np.ndarray
    tx_len: np.ndarray
    cluster_off: np.ndarray = field(default=None)
    member_weight: np.ndarray = field(default=None)

    @property
    def n_cls(self):
        return len(self.cls_off) - 1

    @property
    def nnz(self):
        return int(self.cls_off[-1]) if len(self.cls_off) else 0

    @classmethod
    def from_lists(cls, n_tx, members, counts, tx_len):
        off = np.zeros(len(members) + 1, dtype=np.uint32)
        off[1:] = np.cumsum([len(m) for m in members]) if members else []
        ref = (np.concatenate([np.asarray(m, dtype=np.uint32) for m in members])
               if members and off[-1] else np.zeros(0, dtype=np.uint32))
        return cls(n_tx, off, ref, as_f64(counts), as_u32(tx_len))

    def tiling(self):
        """Component tiling of the table (host only): dict(n_tiles, max_tx,
        max_cls); n_tiles == 0 when a component is too large to tile."""
        c = self.to_c()
        n = _lib.C.c_uint32()
        mt = _lib.C.c_uint32()
        mc = _lib.C.c_uint32()
        check(_lib.load().hem_classes_tiling(c, n, mt, mc))
        return dict(n_tiles=n.value, max_tx=mt.value, max_cls=mc.value)

    def to_c(self):
        self.cls_off = as_u32(self.cls_off)
        self.cls_ref = as_u32(self.cls_ref)
        self.cls_count = as_f64(self.cls_count)
        self.tx_len = as_u32(self.tx_len)
        w = None
        if self.member_weight is not None:
            self.member_weight = as_f64(self.member_weight)
            w = f64p(self.member_weight)
        return _lib.HemClasses(self.n_tx, self.n_cls, u32p(self.cls_off),
                               u32p(self.cls_ref), f64p(self.cls_count),
                               u32p(self.tx_len), w)


def generate(n_tx, n_cls_all, n_frag, seed=0, **overrides):
    """Seeded synthetic class table in hera's layout (include/hera_em_gen.h).

    Returns (Classes, unique_counts).
    """
    lib = _lib.load()
    g = _lib.HemGenParams()
    lib.hem_gen_default(g, n_tx, n_cls_all, n_frag, seed)
    for k, v in overrides.items():
        setattr(g, k, v)
    out = _lib.C.POINTER(_lib.HemGenProblem)()
    check(lib.hem_generate(g, out))
    try:
        p = out.contents
        C, T, Z = p.n_cls, p.n_tx, p.nnz
        cls = Classes(
            T,
            np.ctypeslib.as_array(p.cls_off, (C + 1,)).copy(),
            np.ctypeslib.as_array(p.cls_ref, (max(Z, 1),))[:Z].copy(),
            np.ctypeslib.as_array(p.cls_count, (max(C, 1),))[:C].copy(),
            np.ctypeslib.as_array(p.tx_len, (T,)).copy(),
            np.ctypeslib.as_array(p.cluster_off, (p.n_clusters + 1,)).copy())
        unique = np.ctypeslib.as_array(p.unique, (T,)).copy()
    finally:
        lib.hem_gen_free(out)
    return cls, unique


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth, read + write GB/s
    (hem_copy_bandwidth)."""
    out = _lib.C.c_double()
    check(_lib.load().hem_copy_bandwidth(nbytes, reps, out))
    return out.value


SUM_PLAIN = 0
SUM_NORM_SD = 1
SUM_SQUARES = 2
SUM_WINDOWS = 16   # or'ed into mode: the window method of the exact sums


def exact_sums(rows, mode=SUM_PLAIN):
    """Left-to-right sums (or Norm, or sums of squares) of each row of a
    2-D FP64 array on the device, bitwise the sequential loop
    (hem_exact_sums).  Returns (out, reruns)."""
    x = as_f64(rows)
    if x.ndim == 1:
        x = x[None, :]
    out = np.empty(x.shape[0])
    rr = _lib.C.c_uint32()
    check(_lib.load().hem_exact_sums(f64p(x), x.shape[1], x.shape[0], mode,
                                     f64p(out), rr))
    return out, rr.value


def name_blocks(names, width=None):
    """Names as hera stores them (REF_INF->ref_name, GENE_MAP->gene_name):
    fixed-width blocks, each name NUL-terminated inside its block.
    Returns (bytes, width)."""
    enc = [n.encode() if isinstance(n, str) else bytes(n) for n in names]
    w = width or (max((len(x) for x in enc), default=0) + 1)
    buf = bytearray(w * len(enc))
    for i, x in enumerate(enc):
        if len(x) >= w or b"\0" in x:
            raise ValueError("name %r does not fit a %d-byte block" % (x, w))
        buf[i * w:i * w + len(x)] = x
    return bytes(buf), w


def write_abundance(path, names, unique, length, eff_len, est, tpm,
                    name_len=None):
    """hera's abundance.tsv (hem_write_abundance)."""
    blk, w = name_blocks(names, name_len)
    n = len(names)
    u, ln = as_u32(unique), as_u32(length)
    e, es, tp = as_f64(eff_len), as_f64(est), as_f64(tpm)
    for a in (u, ln, e, es, tp):
        if a.shape != (n,):
            raise ValueError("arrays must have one entry per name")
    check(_lib.load().hem_write_abundance(
        os.fsencode(path), n, blk, w, u32p(u), u32p(ln), f64p(e),
        f64p(es), f64p(tp)))


def write_gene_abundance(path, names, gene, gene_names, est, tpm,
                         name_len=None, gene_len=None):
    """hera's abundance.gene.tsv (hem_write_gene_abundance): names and gene
    names are read from their 17th byte, as the reference does."""
    blk, w = name_blocks(names, name_len)
    gblk, gw = name_blocks(gene_names, gene_len)
    n = len(names)
    g, es, tp = as_u32(gene), as_f64(est), as_f64(tpm)
    for a in (g, es, tp):
        if a.shape != (n,):
            raise ValueError("arrays must have one entry per name")
    check(_lib.load().hem_write_gene_abundance(
        os.fsencode(path), n, blk, w, u32p(g), gblk, gw, len(gene_names),
        f64p(es), f64p(tp)))


def write_result(out_dir, prefix, names, unique, length, eff_len, est, tpm,
                 gene=None, gene_names=None, name_len=None, gene_len=None):
    """Both tables under hera's file names (hem_write_result)."""
    blk, w = name_blocks(names, name_len)
    n = len(names)
    u, ln = as_u32(unique), as_u32(length)
    e, es, tp = as_f64(eff_len), as_f64(est), as_f64(tpm)
    g = gblk = None
    gw = ng = 0
    if gene is not None:
        g = as_u32(gene)
        gblk, gw = name_blocks(gene_names, gene_len)
        ng = len(gene_names)
    check(_lib.load().hem_write_result(
        os.fsencode(out_dir), prefix.encode(), n, blk, w, u32p(u),
        u32p(ln), f64p(e), u32p(g) if g is not None else None, gblk, gw,
        ng, f64p(es), f64p(tp)))


def opts(**kw):
    o = _lib.HemOpts()
    _lib.load().hem_opts_default(o)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def tile_config(config=None):
    """A hem_tile_config from a dict (missing fields: the defaults)."""
    g = _lib.HemTileConfig()
    _lib.load().hem_tile_config_default(g)
    for k, v in (config or {}).items():
        if not hasattr(g, k):
            raise KeyError(k)
        setattr(g, k, int(v))
    return g


def autotune(classes, unique, n_rep=0, iters=20, repeats=3, cache=None):
    """Tile configuration for a class table, timed on the device
    (hem_autotune).  Returns (config dict, report dict)."""
    u = as_u32(unique)
    assert u.shape == (classes.n_tx,)
    o = _lib.HemTuneOpts(n_rep, iters, repeats,
                         os.fsencode(cache) if cache else None)
    g = _lib.HemTileConfig()
    r = _lib.HemTuneReport()
    c = classes.to_c()
    check(_lib.load().hem_autotune(c, u32p(u), o, g, r))
    cfg = {k: getattr(g, k) for k, _ in _lib.HemTileConfig._fields_}
    rep = {k: getattr(r, k) for k, _ in _lib.HemTuneReport._fields_}
    rep["key"] = r.key.decode()
    return cfg, rep


def tune_key(n_tx, n_cls, nnz, n_rep, arch):
    buf = _lib.C.create_string_buffer(160)
    check(_lib.load().hem_tune_key(n_tx, n_cls, nnz, n_rep, arch.encode(),
                                   buf, 160))
    return buf.value.decode()


def tune_cache_lookup(path, key):
    g = _lib.HemTileConfig()
    check(_lib.load().hem_tune_cache_lookup(os.fsencode(path), key.encode(),
                                            g))
    return {k: getattr(g, k) for k, _ in _lib.HemTileConfig._fields_}


def tune_cache_store(path, key, config):
    check(_lib.load().hem_tune_cache_store(os.fsencode(path), key.encode(),
                                           tile_config(config)))


class Problem:
    """A device-resident class table (hem_problem)."""

    def __init__(self, classes, config=None):
        """config: a tile configuration (dict with the fields of
        hem_tile_config, e.g. from autotune(); missing fields keep their
        defaults), or None for the defaults."""
        lib = _lib.load()
        self.classes = classes
        self._c = classes.to_c()
        self._h = _lib.C.c_void_p()
        if config is None:
            check(lib.hem_problem_create(self._c, self._h))
        else:
            check(lib.hem_problem_create_config(self._c, tile_config(config),
                                                self._h))
        self.n_tx = classes.n_tx
        self.n_samples = 1

    def config(self):
        """The tile configuration of the handle (hem_problem_get_config)."""
        g = _lib.HemTileConfig()
        check(_lib.load().hem_problem_get_config(self._h, g))
        return {k: getattr(g, k) for k, _ in _lib.HemTileConfig._fields_}

    @classmethod
    def from_samples(cls, samples):
        """Ragged multi-sample handle (hem_samples_create): the samples'
        tables side by side; batched solves and bootstraps run every
        sample's replicates together.  Count and output rows span all
        samples' transcripts (sample q at columns tx0[q]:tx0[q+1]); seeds and
        iteration counts are per (sample, replicate)."""
        self = cls.__new__(cls)
        lib = _lib.load()
        arr = (_lib.HemClasses * len(samples))(*[c.to_c() for c in samples])
        self._samples = samples
        self._h = _lib.C.c_void_p()
        check(lib.hem_samples_create(arr, len(samples), self._h))
        self.n_samples = len(samples)
        self.tx0 = np.concatenate([[0], np.cumsum([c.n_tx for c in samples])])
        self.cls0 = np.concatenate([[0], np.cumsum([c.n_cls for c in samples])])
        self.n_tx = int(self.tx0[-1])
        off = [np.zeros(1, np.uint32)]
        base = 0
        for c in samples:
            off.append(np.asarray(c.cls_off[1:], np.uint64) + base)
            base += int(c.cls_off[-1]) if c.n_cls else 0
        self.classes = Classes(
            self.n_tx, np.concatenate(off).astype(np.uint32),
            np.concatenate([np.asarray(c.cls_ref, np.uint32) + np.uint32(t)
                            for c, t in zip(samples, self.tx0[:-1])]),
            np.concatenate([c.cls_count for c in samples]),
            np.concatenate([c.tx_len for c in samples]))
        return self

    def close(self):
        if self._h:
            _lib.load().hem_problem_destroy(self._h)
            self._h = _lib.C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_counts(self, cls_count):
        """Replaces the class counts (the structure stays on the device)."""
        cc = as_f64(cls_count)
        assert cc.shape == (self.classes.n_cls,)
        check(_lib.load().hem_problem_set_counts(self._h, f64p(cc)))
        self._counts = cc

    def info(self):
        T = _lib.C.c_uint32()
        Cn = _lib.C.c_uint32()
        Z = _lib.C.c_uint64()
        B = _lib.C.c_uint64()
        check(_lib.load().hem_problem_info(self._h, T, Cn, Z, B))
        return dict(n_tx=T.value, n_cls=Cn.value, nnz=Z.value,
                    device_bytes=B.value)

    def solve(self, unique, **kw):
        """SQUAREM EM solve; returns dict(tpm, est, lam, stats)."""
        u = as_u32(unique)
        assert u.shape == (self.n_tx,)
        o = opts(**kw)
        tpm = np.empty(self.n_tx)
        est = np.empty(self.n_tx)
        lam = np.empty(self.n_tx)
        st = _lib.HemStats()
        check(_lib.load().hem_solve(self._h, u32p(u), o, f64p(tpm),
                                    f64p(est), f64p(lam), st))
        stats = {k: getattr(st, k) for k, _ in _lib.HemStats._fields_}
        return dict(tpm=tpm, est=est, lam=lam, stats=stats)

    def profile(self, unique, **kw):
        """Per-kernel device time of one solve (hem_profile_solve).

        Returns {kind: dict(launches, ms_total, ms_avg, bytes_per_launch)}
        plus 'iters'."""
        u = as_u32(unique)
        o = opts(**kw)
        pr = _lib.HemKernelProfile()
        check(_lib.load().hem_profile_solve(self._h, u32p(u), o, pr))
        out = {"iters": pr.iters}
        for i, name in enumerate(_lib.KERNEL_KIND_NAMES):
            n = pr.launches[i]
            out[name] = dict(launches=n, ms_total=pr.ms[i],
                             ms_avg=pr.ms[i] / n if n else 0.0,
                             bytes_per_launch=pr.bytes[i])
        return out

    def tiles(self):
        n = _lib.C.c_uint32()
        mt = _lib.C.c_uint32()
        mc = _lib.C.c_uint32()
        check(_lib.load().hem_problem_tiles(self._h, n, mt, mc))
        return dict(n_tiles=n.value, max_tx=mt.value, max_cls=mc.value)

    @staticmethod
    def _bstats(st):
        return {k: getattr(st, k) for k, _ in _lib.HemBatchStats._fields_}

    def solve_batch(self, unique, cls_count=None, **kw):
        """R solves sharing this table: unique (R, T) u32, cls_count (R, C)
        float or None.  Returns dict(est, tpm, iters, stats)."""
        u = as_u32(unique)
        R = u.shape[0]
        assert u.shape == (R, self.n_tx)
        cc = None
        if cls_count is not None:
            cc = as_f64(cls_count)
            assert cc.shape == (R, self.classes.n_cls)
        o = opts(**kw)
        est = np.empty((R, self.n_tx))
        tpm = np.empty((R, self.n_tx))
        it = np.empty(self.n_samples * R, dtype=np.uint32)
        st = _lib.HemBatchStats()
        check(_lib.load().hem_solve_batch(
            self._h, R, u32p(u), f64p(cc) if cc is not None else None, o,
            f64p(est), f64p(tpm), u32p(it), st))
        if self.n_samples > 1:
            it = it.reshape(self.n_samples, R)
        return dict(est=est, tpm=tpm, iters=it, stats=self._bstats(st))

    def profile_batch(self, unique, cls_count=None, **kw):
        """Per-kernel device time and algorithmic bytes of the batched
        solves of solve_batch (hem_profile_batch).  Returns {kind:
        dict(launches, ms_total, bytes_total)} plus iters, rep_iters,
        chunk, n_tiles, sum_reruns."""
        u = as_u32(unique)
        R = u.shape[0]
        assert u.shape == (R, self.n_tx)
        cc = None
        if cls_count is not None:
            cc = as_f64(cls_count)
            assert cc.shape == (R, self.classes.n_cls)
        o = opts(**kw)
        pr = _lib.HemBatchProfile()
        check(_lib.load().hem_profile_batch(
            self._h, R, u32p(u), f64p(cc) if cc is not None else None, o,
            pr))
        out = dict(iters=pr.iters, rep_iters=pr.rep_iters, chunk=pr.chunk,
                   n_tiles=pr.n_tiles, sum_reruns=pr.sum_reruns)
        for i, name in enumerate(_lib.BATCH_KIND_NAMES):
            out[name] = dict(launches=pr.launches[i], ms_total=pr.ms[i],
                             bytes_total=pr.bytes[i])
        return out

    def bootstrap(self, unique, seeds, **kw):
        """hera's -b bootstrap, multinomial over all fragments; one seed per
        replicate.  Returns dict(est, tpm, iters, stats)."""
        u = as_u32(unique)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        R = sd.shape[-1]
        assert sd.size == self.n_samples * R
        o = opts(**kw)
        est = np.empty((R, self.n_tx))
        tpm = np.empty((R, self.n_tx))
        it = np.empty(self.n_samples * R, dtype=np.uint32)
        st = _lib.HemBatchStats()
        check(_lib.load().hem_bootstrap(
            self._h, u32p(u), R, sd.ctypes.data_as(_lib.C.POINTER(_lib.C.c_uint64)),
            o, f64p(est), f64p(tpm), u32p(it), st))
        if self.n_samples > 1:
            it = it.reshape(self.n_samples, R)
        return dict(est=est, tpm=tpm, iters=it, stats=self._bstats(st))

    def bootstrap_counts(self, unique, seeds):
        """The resampled (unique, class) counts hem_bootstrap would solve."""
        u = as_u32(unique)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        R = sd.shape[-1]
        assert sd.size == self.n_samples * R
        uo = np.empty((R, self.n_tx), dtype=np.uint32)
        co = np.empty((R, self.classes.n_cls))
        check(_lib.load().hem_bootstrap_counts(
            self._h, u32p(u), R,
            sd.ctypes.data_as(_lib.C.POINTER(_lib.C.c_uint64)), u32p(uo),
            f64p(co)))
        return uo, co

    def bootstrap_hera(self, unique, n_rep, seed=0, **kw):
        """Bootstrap with hera's own sampler (unique reads redrawn with
        glibc rand_r, class counts kept); replicate r continues the rand_r
        state of replicate r - 1 from seed, as hera does on one thread.
        Returns dict(est, tpm, iters, stats)."""
        u = as_u32(unique)
        o = opts(**kw)
        est = np.empty((n_rep, self.n_tx))
        tpm = np.empty((n_rep, self.n_tx))
        it = np.empty(n_rep, dtype=np.uint32)
        st = _lib.HemBatchStats()
        check(_lib.load().hem_bootstrap_hera(
            self._h, u32p(u), n_rep, seed, o, f64p(est), f64p(tpm),
            u32p(it), st))
        return dict(est=est, tpm=tpm, iters=it, stats=self._bstats(st))

    def bootstrap_hera_counts(self, unique, n_rep, seed=0):
        """The unique counts bootstrap_hera would solve, [n_rep, n_tx]."""
        u = as_u32(unique)
        uo = np.empty((n_rep, self.n_tx), dtype=np.uint32)
        check(_lib.load().hem_bootstrap_hera_counts(
            self._h, u32p(u), n_rep, seed, u32p(uo)))
        return uo

    def solve_hera_em(self, noise_tx, p0=None, **kw):
        """hera-EM's simplified-model SQUAREM with locking and filtering
        (hem_solve_hera_em); noise_tx is the noise component's transcript
        (>= n_tx: none).  Returns (p, stats dict)."""
        lib = _lib.load()
        o = _lib.HemHeraEmOpts()
        lib.hem_hera_em_opts_default(o)
        o.noise_tx = noise_tx
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        ini = as_f64(p0) if p0 is not None else None
        out = np.empty(self.n_tx)
        st = _lib.HemHeraEmStats()
        check(lib.hem_solve_hera_em(self._h,
                                    f64p(ini) if ini is not None else None,
                                    o, f64p(out), st))
        return out, {k: getattr(st, k) for k, _ in
                     _lib.HemHeraEmStats._fields_}

    def loglik(self, unique, lam):
        u = as_u32(unique)
        lam = as_f64(lam)
        out = _lib.C.c_double()
        check(_lib.load().hem_loglik(self._h, u32p(u), f64p(lam), out))
        return out.value

    def grad(self, unique, theta, log_efflen, precision=PREC_FP64):
        """Log-likelihood and its gradient in the softmax logits theta and
        the log effective lengths (include/hera_em.h, hem_grad).  Returns
        dict(loglik, n_frag, g_theta, g_log_efflen, ms, ms_adjoint)."""
        u = as_u32(unique)
        th = as_f64(theta)
        el = as_f64(log_efflen)
        gt = np.empty(len(th))
        ge = np.empty(len(th))
        st = _lib.HemGradStats()
        check(_lib.load().hem_grad(self._h, u32p(u), f64p(th), f64p(el),
                                   precision, f64p(gt), f64p(ge), st))
        return dict(loglik=st.loglik, n_frag=st.n_frag, g_theta=gt,
                    g_log_efflen=ge, ms=st.ms, ms_adjoint=st.ms_adjoint)

    def grad_samples(self, unique, theta, log_efflen, precision=PREC_FP64):
        """Every sample's gradient in one pass (hem_grad_samples): the
        softmax within each sample.  Returns dict(loglik[ns], n_frag[ns],
        g_theta, g_log_efflen, ms, ms_adjoint)."""
        u = as_u32(unique)
        th = as_f64(theta)
        el = as_f64(log_efflen)
        assert u.shape == th.shape == el.shape == (self.n_tx,)
        gt = np.empty(self.n_tx)
        ge = np.empty(self.n_tx)
        L = np.empty(self.n_samples)
        N = np.empty(self.n_samples)
        st = _lib.HemGradStats()
        check(_lib.load().hem_grad_samples(self._h, u32p(u), f64p(th),
                                           f64p(el), precision, f64p(gt),
                                           f64p(ge), f64p(L), f64p(N), st))
        return dict(loglik=L, n_frag=N, g_theta=gt, g_log_efflen=ge,
                    ms=st.ms, ms_adjoint=st.ms_adjoint)

    def hvp(self, unique, theta, log_efflen, v):
        """Hessian-vector product in theta (hem_hvp)."""
        u = as_u32(unique)
        th = as_f64(theta)
        el = as_f64(log_efflen)
        v = as_f64(v)
        hv = np.empty(len(th))
        check(_lib.load().hem_hvp(self._h, u32p(u), f64p(th), f64p(el),
                                  f64p(v), f64p(hv)))
        return hv

    def stderr(self, unique, theta, log_efflen, top_k=1000, cg_steps=50,
               damping=0.0):
        """Standard errors of the top_k abundances by batched CG
        (hem_stderr).  Returns dict(idx, se, resid, stats)."""
        u = as_u32(unique)
        th = as_f64(theta)
        el = as_f64(log_efflen)
        idx = np.empty(top_k, dtype=np.uint32)
        se = np.empty(top_k)
        rs = np.empty(top_k)
        st = _lib.HemSeStats()
        check(_lib.load().hem_stderr(self._h, u32p(u), f64p(th), f64p(el),
                                     top_k, cg_steps, damping, u32p(idx),
                                     f64p(se), f64p(rs), st))
        return dict(idx=idx, se=se, resid=rs,
                    stats={k: getattr(st, k) for k, _ in
                           _lib.HemSeStats._fields_})
