"""The tile configuration and the autotuner's cache (host side; no device).

The autotuner itself times candidates on the device (tests/test_gpu_tune.py);
here: the default configuration, the validation of a configuration before
any device work, the cache key and the cache file's round trip, last-line
rule, comments and malformed lines.
"""
import pytest

import hera_amd as H
from hera_amd import _lib
from conftest import tiny_problem


def test_default_config():
    g = H.tile_config()
    assert (g.ptile_tx, g.ptile_cls, g.ptile_nnz) == (64, 320, 1024)
    assert (g.tile_tx, g.tile_cls, g.tile_nnz, g.chunk) == (24, 120, 576, 0)


@pytest.mark.parametrize("bad", [dict(ptile_tx=0), dict(tile_nnz=0),
                                 dict(ptile_tx=65), dict(tile_cls=321),
                                 dict(chunk=3), dict(chunk=64)])
def test_invalid_config_rejected_before_device(bad):
    cls, _ = tiny_problem(1)
    with pytest.raises(H.HemError) as ei:
        H.Problem(cls, config=bad)
    assert ei.value.code == _lib.HEM_ERR_INVALID


def test_unknown_config_field():
    with pytest.raises(KeyError):
        H.tile_config(dict(tiles=3))


def test_key():
    k = H.tune_key(200000, 1300010, 4561171, 100, "gfx950:sramecc+:xnack-")
    assert k == "T=200000,C=1300010,Z=4561171,R=100,arch=gfx950:sramecc+:xnack-"
    with pytest.raises(H.HemError):
        H.tune_key(1, 1, 1, 0, "gfx 950")
    buf = _lib.C.create_string_buffer(8)
    assert _lib.load().hem_tune_key(1, 2, 3, 4, b"gfx950", buf, 8) == \
        _lib.HEM_ERR_RANGE


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.txt")
    k1 = H.tune_key(10, 20, 30, 0, "gfx950")
    k2 = H.tune_key(10, 20, 30, 8, "gfx950")
    with pytest.raises(H.HemError):           # no file yet
        H.tune_cache_lookup(path, k1)
    a = dict(ptile_tx=48, ptile_cls=240, ptile_nnz=768, chunk=8)
    b = dict(tile_tx=16, tile_cls=80, tile_nnz=384, chunk=32)
    H.tune_cache_store(path, k1, a)
    H.tune_cache_store(path, k2, b)
    ga = H.tune_cache_lookup(path, k1)
    gb = H.tune_cache_lookup(path, k2)
    want = H.tile_config(a)
    assert ga == {k: getattr(want, k) for k, _ in _lib.HemTileConfig._fields_}
    assert gb["tile_tx"] == 16 and gb["chunk"] == 32 and gb["ptile_tx"] == 64
    # the last line of a key wins; comments and malformed lines are skipped
    with open(path, "a") as f:
        f.write("# a comment %s 1 1 1 1 1 1 1\n" % k1)
        f.write("%s 7 8\n" % k1)
        f.write("%s 32 160 512 24 120 576 16\n" % k1)
    assert H.tune_cache_lookup(path, k1)["ptile_tx"] == 32
    assert H.tune_cache_lookup(path, k1)["chunk"] == 16
    with pytest.raises(H.HemError) as ei:
        H.tune_cache_lookup(path, H.tune_key(1, 1, 1, 0, "gfx950"))
    assert ei.value.code == _lib.HEM_ERR_INVALID
