"""ctypes view of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The FP64 CPU restatement of hera's EM (oracle/em_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
always as the checker or the timed CPU baseline, never as a product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


class OracleOpts(C.Structure):
    _fields_ = [("min_round", C.c_uint32), ("max_round", C.c_uint32),
                ("stop_min", C.c_double), ("stop_rel", C.c_double),
                ("step_form", C.c_int), ("clamp_form", C.c_int),
                ("alpha_clamp", C.c_int), ("fallback", C.c_int)]


class OracleHeOpts(C.Structure):
    _fields_ = [("noise_tx", C.c_uint32), ("max_iter", C.c_uint32),
                ("locking", C.c_int), ("filtering", C.c_int),
                ("lock_thr", C.c_double), ("filter_thr", C.c_double),
                ("stop_rel", C.c_double), ("stop_min", C.c_double)]


class OracleHeStats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("converged", C.c_uint32),
                ("n_locked", C.c_uint32), ("removed_members", C.c_uint32),
                ("removed_classes", C.c_uint32)]


class OracleTrace(C.Structure):
    _fields_ = [("cap", C.c_uint32), ("k", C.c_uint32),
                ("alpha", _f64p), ("stop", _u32p), ("last_u", _f64p),
                ("fallback", _u32p)]


_LIB = None


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        lib = C.CDLL(LIB_PATH)
        lib.oracle_norm_sd.restype = C.c_double
        lib.oracle_norm_sd.argtypes = [_f64p, C.c_uint32]
        lib.oracle_estimate_count.restype = C.c_int
        lib.oracle_estimate_count.argtypes = [
            C.c_uint32, C.c_uint32, _u32p, _u32p, _f64p, _u32p, _u32p,
            C.POINTER(OracleOpts), _f64p, _f64p, _f64p, _f64p,
            C.POINTER(OracleTrace)]
        lib.oracle_estimate_count_w.restype = C.c_int
        lib.oracle_estimate_count_w.argtypes = [
            C.c_uint32, C.c_uint32, _u32p, _u32p, _f64p, _f64p, _u32p, _u32p,
            C.POINTER(OracleOpts), _f64p, _f64p, _f64p, _f64p,
            C.POINTER(OracleTrace)]
        lib.oracle_em_map_w.restype = C.c_int
        lib.oracle_em_map_w.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                        _f64p, _f64p, _u32p, _u32p, _f64p,
                                        _f64p, _u32p]
        lib.oracle_write_abundance.restype = C.c_int
        lib.oracle_write_abundance.argtypes = [
            C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, _u32p, _u32p,
            _f64p, _f64p, _f64p]
        lib.oracle_write_gene_abundance.restype = C.c_int
        lib.oracle_write_gene_abundance.argtypes = [
            C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, _u32p,
            C.c_char_p, C.c_uint32, _f64p, _f64p]
        lib.oracle_rand_r.restype = C.c_uint32
        lib.oracle_rand_r.argtypes = [_u32p]
        lib.oracle_hera_sample.restype = C.c_int
        lib.oracle_hera_sample.argtypes = [_u32p, C.c_uint32, _u32p, _u32p]
        lib.oracle_em_map.restype = C.c_int
        lib.oracle_em_map.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                      _f64p, _u32p, _u32p, _f64p, _f64p,
                                      _u32p]
        lib.oracle_loglik.restype = C.c_double
        lib.oracle_loglik.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                      _f64p, _u32p, _u32p, _f64p]
        lib.oracle_grad.restype = C.c_double
        lib.oracle_grad.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                    _f64p, _u32p, _f64p, _f64p, _f64p, _f64p]
        lib.oracle_hvp.restype = None
        lib.oracle_hvp.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                   _f64p, _u32p, _f64p, _f64p, _f64p, _f64p]
        lib.oracle_bootstrap_batch.restype = C.c_int
        lib.oracle_bootstrap_batch.argtypes = [
            C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint32, _u32p,
            _f64p, C.POINTER(OracleOpts), C.c_uint32, _f64p, _u32p]
        lib.oracle_philox.restype = None
        lib.oracle_philox.argtypes = [C.c_uint64, _u32p, _u32p]
        lib.oracle_resample_plan.restype = C.c_uint32
        lib.oracle_resample_plan.argtypes = [_u32p, C.c_uint32, C.c_uint32,
                                             _u32p, _u32p]
        lib.oracle_resample.restype = C.c_int
        lib.oracle_resample.argtypes = [_u32p, C.c_uint32, C.c_uint64, _u32p]
        lib.oracle_bootstrap_resampled.restype = C.c_int
        lib.oracle_bootstrap_resampled.argtypes = [
            C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, _u32p, C.c_uint32,
            C.POINTER(C.c_uint64), C.POINTER(OracleOpts), C.c_uint32, _f64p,
            _u32p]
        lib.oracle_now.restype = C.c_double
        lib.oracle_hera_em.restype = C.c_int
        lib.oracle_hera_em.argtypes = [
            C.c_uint32, C.c_uint32, _u32p, _u32p, _f64p, _f64p,
            C.POINTER(OracleHeOpts), _f64p, _f64p, C.POINTER(OracleHeStats)]
        lib.oracle_dloglik.restype = C.c_double
        lib.oracle_dloglik.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p,
                                       _f64p, _f64p, _u32p, _u32p, _f64p,
                                       _f64p, _f64p]
        lib.oracle_opts_default.argtypes = [C.POINTER(OracleOpts)]
        _LIB = lib
    return _LIB


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p32(a):
    return a.ctypes.data_as(_u32p)


def _p64(a):
    return a.ctypes.data_as(_f64p)


def opts(**kw):
    o = OracleOpts()
    load().oracle_opts_default(o)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def norm_sd(x):
    x = _f64(x)
    return load().oracle_norm_sd(_p64(x), len(x))


def estimate_count(cls, unique, init=None, trace_cap=1000, **kw):
    """Runs the restated estimate_count; returns dict(tpm, est, lam, k,
    alpha, stop)."""
    lib = load()
    T, Cn = cls.n_tx, cls.n_cls
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, ln, u = _f64(cls.cls_count), _u32(cls.tx_len), _u32(unique)
    tpm, est, lam = np.empty(T), np.empty(T), np.empty(T)
    alpha = np.zeros(max(trace_cap, 1))
    stop = np.zeros(max(trace_cap, 1), dtype=np.uint32)
    last_u = np.zeros(T)
    fb = np.zeros(max(trace_cap, 1), dtype=np.uint32)
    tr = OracleTrace(trace_cap, 0, _p64(alpha), _p32(stop), _p64(last_u),
                     _p32(fb))
    ini = _f64(init) if init is not None else None
    w = getattr(cls, "member_weight", None)
    w = _f64(w) if w is not None else None
    rc = lib.oracle_estimate_count_w(T, Cn, _p32(off), _p32(ref), _p64(cnt),
                                     _p64(w) if w is not None else None,
                                     _p32(ln), _p32(u), opts(**kw),
                                     _p64(ini) if ini is not None else None,
                                     _p64(tpm), _p64(est), _p64(lam), tr)
    assert rc == 0
    k = tr.k
    return dict(tpm=tpm, est=est, lam=lam, k=k,
                alpha=alpha[:min(k, trace_cap)].copy(),
                stop=stop[:min(k, trace_cap)].copy(), last_u=last_u,
                fallback=fb[:min(k, trace_cap)].copy())


def dloglik(cls, unique, lam, mu):
    """(loglik(lam) - loglik(mu), scale): member weights in D_c when the
    table has them, summed as per-term differences; scale is the sum of the
    terms' magnitudes."""
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, ln, u = _f64(cls.cls_count), _u32(cls.tx_len), _u32(unique)
    lam, mu = _f64(lam), _f64(mu)
    w = getattr(cls, "member_weight", None)
    w = _f64(w) if w is not None else None
    sc = C.c_double()
    d = load().oracle_dloglik(cls.n_tx, cls.n_cls, _p32(off), _p32(ref),
                              _p64(cnt), _p64(w) if w is not None else None,
                              _p32(ln), _p32(u), _p64(lam), _p64(mu),
                              C.byref(sc))
    return d, sc.value


def em_map(cls, unique, lam):
    lib = load()
    T = cls.n_tx
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, ln, u, lam = (_f64(cls.cls_count), _u32(cls.tx_len), _u32(unique),
                       _f64(lam))
    out = np.empty(T)
    stop = C.c_uint32()
    w = getattr(cls, "member_weight", None)
    w = _f64(w) if w is not None else None
    assert lib.oracle_em_map_w(T, cls.n_cls, _p32(off), _p32(ref), _p64(cnt),
                               _p64(w) if w is not None else None,
                               _p32(ln), _p32(u), _p64(lam), _p64(out),
                               C.byref(stop)) == 0
    return out, stop.value


def loglik(cls, unique, lam):
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, ln, u, lam = (_f64(cls.cls_count), _u32(cls.tx_len), _u32(unique),
                       _f64(lam))
    return load().oracle_loglik(cls.n_tx, cls.n_cls, _p32(off), _p32(ref),
                                _p64(cnt), _p32(ln), _p32(u), _p64(lam))


def grad(cls, unique, theta, ell):
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, u = _f64(cls.cls_count), _u32(unique)
    theta, ell = _f64(theta), _f64(ell)
    gt, ge = np.empty(cls.n_tx), np.empty(cls.n_tx)
    L = load().oracle_grad(cls.n_tx, cls.n_cls, _p32(off), _p32(ref),
                           _p64(cnt), _p32(u), _p64(theta), _p64(ell),
                           _p64(gt), _p64(ge))
    return L, gt, ge


def hvp(cls, unique, theta, ell, vec):
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt, u = _f64(cls.cls_count), _u32(unique)
    theta, ell, vec = _f64(theta), _f64(ell), _f64(vec)
    out = np.empty(cls.n_tx)
    load().oracle_hvp(cls.n_tx, cls.n_cls, _p32(off), _p32(ref), _p64(cnt),
                      _p32(u), _p64(theta), _p64(ell), _p64(vec), _p64(out))
    return out


def bootstrap_batch(cls, counts, cls_counts, n_thread, **kw):
    """Reference join-per-batch thread schedule over R count vectors."""
    R = counts.shape[0]
    off, ref, ln = _u32(cls.cls_off), _u32(cls.cls_ref), _u32(cls.tx_len)
    counts, cls_counts = _u32(counts), _f64(cls_counts)
    est = np.empty((R, cls.n_tx))
    ks = np.empty(R, dtype=np.uint32)
    assert load().oracle_bootstrap_batch(
        cls.n_tx, cls.n_cls, _p32(off), _p32(ref), _p32(ln), R,
        _p32(counts), _p64(cls_counts), opts(**kw), n_thread, _p64(est),
        _p32(ks)) == 0
    return est, ks


def now():
    return load().oracle_now()


def philox(key, ctr):
    c = _u32(ctr)
    out = np.empty(4, dtype=np.uint32)
    load().oracle_philox(key, _p32(c), _p32(out))
    return out


def resample_plan(w):
    w = _u32(w)
    N = int(w.sum(dtype=np.uint64))
    start = np.empty(7002, dtype=np.uint32)
    cw = np.empty(7002, dtype=np.uint32)
    B = load().oracle_resample_plan(_p32(w), len(w), N, _p32(start), _p32(cw))
    return start[:B + 1].copy(), cw[:B + 1].copy()


def resample(w, seed):
    """One replicate of Multinomial(sum w, w / sum w), as the library draws
    it (two-level, Philox + Lemire)."""
    w = _u32(w)
    out = np.empty(len(w), dtype=np.uint32)
    assert load().oracle_resample(_p32(w), len(w), seed, _p32(out)) == 0
    return out


def bootstrap_resampled(cls, unique, seeds, n_thread, **kw):
    """hera's join-per-batch bootstrap schedule; each thread resamples all
    fragments (as resample()) and solves.  Returns (est[R, T], k[R])."""
    off, ref, ln = _u32(cls.cls_off), _u32(cls.cls_ref), _u32(cls.tx_len)
    w = _u32(np.concatenate([np.asarray(unique, dtype=np.uint32),
                             np.asarray(cls.cls_count).astype(np.uint32)]))
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    R = len(sd)
    est = np.empty((R, cls.n_tx))
    ks = np.empty(R, dtype=np.uint32)
    assert load().oracle_bootstrap_resampled(
        cls.n_tx, cls.n_cls, _p32(off), _p32(ref), _p32(ln), _p32(w), R,
        sd.ctypes.data_as(C.POINTER(C.c_uint64)), opts(**kw), n_thread,
        _p64(est), _p32(ks)) == 0
    return est, ks


def write_abundance(path, blocks, width, unique, length, eff_len, est, tpm):
    """src/EM.c:176-185 restated (names as fixed-width NUL-padded blocks)."""
    u = np.ascontiguousarray(unique, np.uint32)
    ln = np.ascontiguousarray(length, np.uint32)
    e, es, tp = (np.ascontiguousarray(a, np.float64)
                 for a in (eff_len, est, tpm))
    if load().oracle_write_abundance(
            os.fsencode(path), len(u), blocks, width,
            u.ctypes.data_as(_u32p), ln.ctypes.data_as(_u32p),
            e.ctypes.data_as(_f64p), es.ctypes.data_as(_f64p),
            tp.ctypes.data_as(_f64p)):
        raise OSError("oracle write failed: %s" % path)


def write_gene_abundance(path, blocks, width, gene, gblocks, gwidth, est,
                         tpm):
    """src/EM.c:187-205 restated."""
    g = np.ascontiguousarray(gene, np.uint32)
    es, tp = (np.ascontiguousarray(a, np.float64) for a in (est, tpm))
    if load().oracle_write_gene_abundance(
            os.fsencode(path), len(g), blocks, width,
            g.ctypes.data_as(_u32p), gblocks, gwidth,
            es.ctypes.data_as(_f64p), tp.ctypes.data_as(_f64p)):
        raise OSError("oracle write failed: %s" % path)


def rand_r_seq(seed, n):
    """n successive values of glibc's rand_r from seed (restated);
    returns (values, final seed)."""
    st = C.c_uint32(seed)
    lib = load()
    out = np.empty(n, np.uint32)
    for i in range(n):
        out[i] = lib.oracle_rand_r(C.byref(st))
    return out, st.value


def hera_sample(unique, n_rep, seed=0):
    """hera's bootstrap_sample (src/EM.c:123-150) for n_rep replicates run
    one after another on one global seed, as hera with one thread."""
    u = np.ascontiguousarray(unique, np.uint32)
    st = C.c_uint32(seed)
    out = np.empty((n_rep, len(u)), np.uint32)
    for r in range(n_rep):
        row = out[r]
        if load().oracle_hera_sample(u.ctypes.data_as(_u32p), len(u),
                                     C.byref(st), row.ctypes.data_as(_u32p)):
            raise MemoryError("oracle sampler")
    return out, st.value


HE_DEFAULTS = dict(max_iter=1000, locking=1, filtering=1, lock_thr=1e-5,
                   filter_thr=1e-6, stop_rel=1e-3, stop_min=1e-7)


def hera_em(cls, noise_tx, p0=None, **kw):
    """hera-EM's simplified-model SQUAREM on a weighted class table (the
    noise component is transcript noise_tx; >= n_tx for none).  Returns
    (p, stats dict)."""
    o = dict(HE_DEFAULTS, **kw)
    opt = OracleHeOpts(noise_tx, o["max_iter"], o["locking"],
                       o["filtering"], o["lock_thr"], o["filter_thr"],
                       o["stop_rel"], o["stop_min"])
    off, ref = _u32(cls.cls_off), _u32(cls.cls_ref)
    cnt = _f64(cls.cls_count)
    w = getattr(cls, "member_weight", None)
    w = _f64(w) if w is not None else None
    out = np.empty(cls.n_tx)
    st = OracleHeStats()
    ini = _f64(p0) if p0 is not None else None
    assert load().oracle_hera_em(cls.n_tx, cls.n_cls, _p32(off), _p32(ref),
                                 _p64(cnt), _p64(w) if w is not None else None,
                                 opt,
                                 _p64(ini) if ini is not None else None,
                                 _p64(out), st) == 0
    return out, {k: getattr(st, k) for k, _ in OracleHeStats._fields_}
