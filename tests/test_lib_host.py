"""Host-side checks of libheraem.so that need no GPU: the C ABI exports
every symbol include/*.h declares, input validation, the generator."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import hera_amd as H
from hera_amd import _lib
from conftest import ROOT


def declared_symbols():
    """Function and data names declared in include/*.h."""
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", text,
                      flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;",
                             text, flags=re.M):
            names.add(m.group(1))
        for m in re.finditer(r"^extern\s+[\w\s\*]+?\b(\w+)\s*;", text,
                             flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = declared_symbols()
    assert {"hem_solve", "estimate_count", "REF_INF", "CLASS"} <= decl
    for name in sorted(decl):
        assert hasattr(lib, name), name
    # and the ctypes table covers every declared function
    assert decl - set(_lib.DATA_SYMBOLS) <= set(_lib.SIGNATURES), \
        decl - set(_lib.SIGNATURES)


def test_opts_defaults_are_the_reference_constants():
    o = H.opts()
    assert (o.min_iter, o.max_iter) == (50, 1000)            # src/EM.c:325-326
    assert (o.stop_min_value, o.stop_rel_change) == (1e-3, 0.01)  # :292,295
    assert o.step == H.STEP_SD and o.clamp == H.CLAMP_ZERO and o.alpha_clamp == 1
    assert o.fallback == 0


@pytest.mark.parametrize("mutate,expect", [
    (lambda c: c.cls_ref.__setitem__(0, c.n_tx), "transcript id"),
    (lambda c: c.tx_len.__setitem__(3, 0), "length 0"),
    (lambda c: c.cls_count.__setitem__(1, -1.0), "negative"),
    (lambda c: c.cls_count.__setitem__(1, np.nan), "not finite"),
    (lambda c: c.cls_off.__setitem__(2, c.cls_off[1] - 1), "monotone"),
])
def test_invalid_tables_rejected_before_device(mutate, expect):
    from conftest import tiny_problem
    cls, _ = tiny_problem(0)
    mutate(cls)
    with pytest.raises(H.HemError) as ei:
        H.Problem(cls)
    assert ei.value.code == _lib.HEM_ERR_INVALID
    assert expect in str(ei.value)


def test_generator_layout_rules(cfg1):
    cls, u = cfg1
    sizes = np.diff(cls.cls_off)
    assert cls.n_tx == 20000
    assert sizes.min() >= 2 and sizes.max() <= 16       # capped by cluster size
    co = cls.cluster_off
    cl_sizes = np.diff(co)
    assert cl_sizes[:-1].min() >= 2 and cl_sizes.max() <= 16
    # members distinct, ascending, inside one cluster
    for c in range(0, cls.n_cls, 997):
        m = cls.cls_ref[cls.cls_off[c]:cls.cls_off[c + 1]]
        assert np.all(np.diff(m.astype(np.int64)) > 0)
        g = np.searchsorted(co, m, side="right") - 1
        assert np.all(g == g[0])
    # integer lengths in [200, 20000]
    assert cls.tx_len.min() >= 200 and cls.tx_len.max() <= 20000
    # ~35% of drawn classes are singletons folded into unique counts; the
    # total is the Zipf-scaled 10M (rounding, min 1)
    total = u.sum() + cls.cls_count.sum()
    assert abs(total - 10_000_000) < 0.02 * 10_000_000
    assert 0.30 < 1 - cls.n_cls / 200000 < 0.40


def test_generator_is_deterministic():
    a, ua = H.generate(3000, 20000, 1_000_000, seed=7)
    b, ub = H.generate(3000, 20000, 1_000_000, seed=7)
    c, uc = H.generate(3000, 20000, 1_000_000, seed=8)
    for x, y in ((a.cls_ref, b.cls_ref), (a.cls_count, b.cls_count),
                 (a.tx_len, b.tx_len), (ua, ub)):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a.tx_len, c.tx_len)


def test_dropin_globals_are_weak_and_writable():
    lib = _lib.load()
    ref = C.c_void_p.in_dll(lib, "REF_INF")
    cls = C.c_void_p.in_dll(lib, "CLASS")
    assert ref.value is None and cls.value is None
