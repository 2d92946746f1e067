"""ctypes bindings of libheraem.so (the C ABI in include/hera_em.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C hera_amd/csrc``).  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheraem.so")

HEM_OK = 0
HEM_ERR_INVALID = 1
HEM_ERR_NOMEM = 2
HEM_ERR_DEVICE = 3
HEM_ERR_NODEVICE = 4
HEM_ERR_RANGE = 5
HEM_ERR_IO = 6

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


class HemClasses(C.Structure):
    _fields_ = [("n_tx", C.c_uint32), ("n_cls", C.c_uint32),
                ("cls_off", _u32p), ("cls_ref", _u32p),
                ("cls_count", _f64p), ("tx_len", _u32p),
                ("member_weight", _f64p)]


class HemOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("step", C.c_int),
                ("clamp", C.c_int), ("alpha_clamp", C.c_int),
                ("norm", C.c_int),
                ("min_iter", C.c_uint32), ("max_iter", C.c_uint32),
                ("stop_min_value", C.c_double),
                ("stop_rel_change", C.c_double),
                ("fallback", C.c_int), ("verbose", C.c_int)]


class HemStats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("maps", C.c_uint32),
                ("converged", C.c_uint32), ("fallbacks", C.c_uint32),
                ("loglik", C.c_double), ("last_alpha", C.c_double),
                ("ms_solve", C.c_double), ("ms_total", C.c_double),
                ("sum_reruns", C.c_uint32)]


HEM_KERNEL_KINDS = 8
KERNEL_KIND_NAMES = ["class_pass", "transcript_pass", "norm_pass1",
                     "norm_pass2", "extrapolate", "iter_end", "tile_map12",
                     "tile_ex3_map12"]


class HemKernelProfile(C.Structure):
    _fields_ = [("iters", C.c_uint32),
                ("launches", C.c_uint32 * HEM_KERNEL_KINDS),
                ("ms", C.c_double * HEM_KERNEL_KINDS),
                ("bytes", C.c_double * HEM_KERNEL_KINDS)]


HEM_BATCH_KINDS = 5
BATCH_KIND_NAMES = ["tile_map12", "norm_pass1", "norm_pass2",
                    "tile_ex3", "iter_end"]


class HemBatchProfile(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("chunk", C.c_uint32),
                ("n_tiles", C.c_uint32), ("sum_reruns", C.c_uint32),
                ("rep_iters", C.c_uint64),
                ("launches", C.c_uint32 * HEM_BATCH_KINDS),
                ("ms", C.c_double * HEM_BATCH_KINDS),
                ("bytes", C.c_double * HEM_BATCH_KINDS)]


class HemBatchStats(C.Structure):
    _fields_ = [("n_rep", C.c_uint32), ("iters_min", C.c_uint32),
                ("iters_max", C.c_uint32), ("n_converged", C.c_uint32),
                ("chunk", C.c_uint32), ("n_tiles", C.c_uint32),
                ("ms_resample", C.c_double), ("ms_solve", C.c_double),
                ("ms_total", C.c_double), ("sum_reruns", C.c_uint32)]


class HemHeraEmOpts(C.Structure):
    _fields_ = [("noise_tx", C.c_uint32), ("max_iter", C.c_uint32),
                ("locking", C.c_int), ("filtering", C.c_int),
                ("lock_threshold", C.c_double),
                ("filter_threshold", C.c_double),
                ("stop_rel_change", C.c_double),
                ("stop_min_value", C.c_double)]


class HemHeraEmStats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("converged", C.c_uint32),
                ("n_locked", C.c_uint32), ("removed_members", C.c_uint32),
                ("removed_classes", C.c_uint32), ("ms", C.c_double)]


class HemGradStats(C.Structure):
    _fields_ = [("loglik", C.c_double), ("n_frag", C.c_double),
                ("ms", C.c_double), ("ms_adjoint", C.c_double)]


class HemSeStats(C.Structure):
    _fields_ = [("top_k", C.c_uint32), ("cg_steps", C.c_uint32),
                ("chunk", C.c_uint32), ("resid_max", C.c_double),
                ("ms_setup", C.c_double), ("ms_cg", C.c_double),
                ("ms_total", C.c_double)]


class HemTileConfig(C.Structure):
    _fields_ = [("ptile_tx", C.c_uint32), ("ptile_cls", C.c_uint32),
                ("ptile_nnz", C.c_uint32), ("tile_tx", C.c_uint32),
                ("tile_cls", C.c_uint32), ("tile_nnz", C.c_uint32),
                ("chunk", C.c_uint32)]


class HemTuneOpts(C.Structure):
    _fields_ = [("n_rep", C.c_uint32), ("iters", C.c_uint32),
                ("repeats", C.c_uint32), ("cache_path", C.c_char_p)]


class HemTuneReport(C.Structure):
    _fields_ = [("from_cache", C.c_int), ("n_candidates", C.c_uint32),
                ("us_plain_default", C.c_double),
                ("us_plain_best", C.c_double),
                ("us_batch_default", C.c_double),
                ("us_batch_best", C.c_double), ("key", C.c_char * 160)]


class HemGenParams(C.Structure):
    _fields_ = [("n_tx", C.c_uint32), ("n_cls_all", C.c_uint32),
                ("n_frag", C.c_uint64), ("cluster_min", C.c_uint32),
                ("cluster_max", C.c_uint32), ("len_mu", C.c_double),
                ("len_sigma", C.c_double), ("len_min", C.c_uint32),
                ("len_max", C.c_uint32), ("geom_p", C.c_double),
                ("cls_max", C.c_uint32), ("zipf_s", C.c_double),
                ("seed", C.c_uint64)]


class HemGenProblem(C.Structure):
    _fields_ = [("n_tx", C.c_uint32), ("n_cls", C.c_uint32),
                ("nnz", C.c_uint64), ("cls_off", _u32p), ("cls_ref", _u32p),
                ("cls_count", _f64p), ("tx_len", _u32p), ("unique", _u32p),
                ("n_clusters", C.c_uint32), ("cluster_off", _u32p),
                ("frag_total", C.c_uint64)]


# Every symbol include/*.h declares, with its ctypes signature.
SIGNATURES = {
    "hem_last_error": (C.c_char_p, []),
    "hem_opts_default": (None, [C.POINTER(HemOpts)]),
    "hem_problem_create": (C.c_int, [C.POINTER(HemClasses), C.POINTER(C.c_void_p)]),
    "hem_problem_destroy": (C.c_int, [C.c_void_p]),
    "hem_tile_config_default": (None, [C.POINTER(HemTileConfig)]),
    "hem_problem_create_config": (C.c_int, [C.POINTER(HemClasses),
                                            C.POINTER(HemTileConfig),
                                            C.POINTER(C.c_void_p)]),
    "hem_problem_get_config": (C.c_int, [C.c_void_p,
                                         C.POINTER(HemTileConfig)]),
    "hem_autotune": (C.c_int, [C.POINTER(HemClasses), _u32p,
                               C.POINTER(HemTuneOpts),
                               C.POINTER(HemTileConfig),
                               C.POINTER(HemTuneReport)]),
    "hem_tune_key": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64,
                               C.c_uint32, C.c_char_p, C.c_char_p,
                               C.c_size_t]),
    "hem_tune_cache_lookup": (C.c_int, [C.c_char_p, C.c_char_p,
                                        C.POINTER(HemTileConfig)]),
    "hem_tune_cache_store": (C.c_int, [C.c_char_p, C.c_char_p,
                                       C.POINTER(HemTileConfig)]),
    "hem_samples_create": (C.c_int, [C.POINTER(HemClasses), C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "hem_problem_set_counts": (C.c_int, [C.c_void_p, _f64p]),
    "hem_problem_info": (C.c_int, [C.c_void_p, _u32p, _u32p,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "hem_solve": (C.c_int, [C.c_void_p, _u32p, C.POINTER(HemOpts), _f64p,
                            _f64p, _f64p, C.POINTER(HemStats)]),
    "hem_loglik": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p]),
    "hem_write_abundance": (C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p,
                                      C.c_uint32, _u32p, _u32p, _f64p,
                                      _f64p, _f64p]),
    "hem_write_gene_abundance": (C.c_int, [C.c_char_p, C.c_uint32,
                                           C.c_char_p, C.c_uint32, _u32p,
                                           C.c_char_p, C.c_uint32,
                                           C.c_uint32, _f64p, _f64p]),
    "hem_write_result": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint32,
                                   C.c_char_p, C.c_uint32, _u32p, _u32p,
                                   _f64p, _u32p, C.c_char_p, C.c_uint32,
                                   C.c_uint32, _f64p, _f64p]),
    "hem_exact_sums": (C.c_int, [_f64p, C.c_uint32, C.c_uint32, C.c_int,
                                 _f64p, _u32p]),
    "hem_exact_sums_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32,
                                         C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    "hem_profile_solve": (C.c_int, [C.c_void_p, _u32p, C.POINTER(HemOpts),
                                    C.POINTER(HemKernelProfile)]),
    "hem_profile_batch": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _f64p,
                                    C.POINTER(HemOpts),
                                    C.POINTER(HemBatchProfile)]),
    "hem_problem_tiles": (C.c_int, [C.c_void_p, _u32p, _u32p, _u32p]),
    "hem_classes_tiling": (C.c_int, [C.POINTER(HemClasses), _u32p, _u32p,
                                     _u32p]),
    "hem_solve_batch": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _f64p,
                                  C.POINTER(HemOpts), _f64p, _f64p, _u32p,
                                  C.POINTER(HemBatchStats)]),
    "hem_bootstrap": (C.c_int, [C.c_void_p, _u32p, C.c_uint32,
                                C.POINTER(C.c_uint64), C.POINTER(HemOpts),
                                _f64p, _f64p, _u32p,
                                C.POINTER(HemBatchStats)]),
    "hem_bootstrap_counts": (C.c_int, [C.c_void_p, _u32p, C.c_uint32,
                                       C.POINTER(C.c_uint64), _u32p, _f64p]),
    "hem_bootstrap_hera": (C.c_int, [C.c_void_p, _u32p, C.c_uint32,
                                     C.c_uint32, C.POINTER(HemOpts), _f64p,
                                     _f64p, _u32p,
                                     C.POINTER(HemBatchStats)]),
    "hem_bootstrap_hera_counts": (C.c_int, [C.c_void_p, _u32p, C.c_uint32,
                                            C.c_uint32, _u32p]),
    "hem_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_int,
                                     C.POINTER(C.c_double)]),
    "hem_hera_em_opts_default": (None, [C.POINTER(HemHeraEmOpts)]),
    "hem_solve_hera_em": (C.c_int, [C.c_void_p, _f64p,
                                    C.POINTER(HemHeraEmOpts), _f64p,
                                    C.POINTER(HemHeraEmStats)]),
    "hem_grad": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p, C.c_int, _f64p,
                           _f64p, C.POINTER(HemGradStats)]),
    "hem_grad_samples": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p,
                                   C.c_int, _f64p, _f64p, _f64p, _f64p,
                                   C.POINTER(HemGradStats)]),
    "hem_hvp": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p, _f64p, _f64p]),
    "hem_stderr": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p, C.c_uint32,
                             C.c_uint32, C.c_double, _u32p, _f64p, _f64p,
                             C.POINTER(HemSeStats)]),
    "hem_gen_default": (None, [C.POINTER(HemGenParams), C.c_uint32,
                               C.c_uint32, C.c_uint64, C.c_uint64]),
    "hem_generate": (C.c_int, [C.POINTER(HemGenParams),
                               C.POINTER(C.POINTER(HemGenProblem))]),
    "hem_gen_free": (None, [C.POINTER(HemGenProblem)]),
    "estimate_count": (C.c_void_p, [_u32p, C.c_uint]),
    "destroy_emVal": (None, [C.c_void_p]),
    "hem_dropin_release": (C.c_int, []),
}

# Global variables the drop-in header declares.
DATA_SYMBOLS = ["REF_INF", "CLASS"]

_LIB = None


def load():
    """Load libheraem.so once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libheraem.so is not built (expected at %s); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`make -C hera_amd/csrc`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class HemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hera_em error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != HEM_OK:
        raise HemError(rc, load().hem_last_error().decode())
    return rc


def u32p(a):
    return a.ctypes.data_as(_u32p)


def f64p(a):
    return a.ctypes.data_as(_f64p)


def as_u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
