"""Autotuner host logic (no GPU): cache keys, candidate lists, the JSON
table, and the choice of the fastest candidate with a stand-in timer."""
import json

import pytest

import bigdft_hip
from bigdft_hip import autotune
from bigdft_hip.descriptors import dense_region


class FakeCtx:
    def __init__(self):
        self.opts = {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0,
                     "tile_order": 0, "persist": 0}
        self.history = []

    def set_option(self, k, v):
        assert k in bigdft_hip.OPTIONS
        self.opts[k] = v
        self.history.append((k, v))


class FakeLr:
    def __init__(self, region):
        self.region = region


def test_shape_keys():
    a = dense_region("P", (47, 47, 47), (0.35,) * 3)
    b = dense_region("F", (20, 31, 17), (0.4,) * 3)
    c = dense_region("F", (25, 18, 19), (0.4,) * 3)
    assert autotune.shape_key([a] * 4, bigdft_hip.PREC_F16X3) == "P:47x47x47:f16x3"
    k = autotune.shape_key([b, c], bigdft_hip.PREC_FP64)
    assert k.startswith("F:max25x31x19:mixed") and k.endswith(":fp64")
    assert autotune.shape_key([a, b], bigdft_hip.PREC_FP32).startswith("mixed:max47x47x47")
    # batches with the same largest extents but other boxes get other keys
    d = dense_region("F", (25, 31, 19), (0.4,) * 3)
    assert autotune.shape_key([b, c, d], bigdft_hip.PREC_FP64) != k
    assert autotune.shape_key([c, b], bigdft_hip.PREC_FP64) == k
    # the device part keeps tables of other GPUs or library versions apart
    assert autotune.shape_key([a], bigdft_hip.PREC_FP64, "gfx950/lds163840/v3") == \
        "gfx950/lds163840/v3|P:47x47x47:fp64"


def test_candidates():
    assert len(autotune.candidates(bigdft_hip.PREC_FP64)) == 2
    c16 = autotune.candidates(bigdft_hip.PREC_F16X3)
    assert len(c16) == 8 and {x["zpot"] for x in c16} == {0}
    c = autotune.candidates(bigdft_hip.PREC_BF16X4)
    assert len(c) == 14 and c[0] == {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0,
                                     "chunk_mb": 0, "tile_order": 0, "persist": 0}
    assert {x["zgroup"] for x in c} == {0, 2, 3} and {x["zpot"] for x in c} == {0, 1}
    assert {x["bwd"] for x in c} == {0, 1, 2}
    assert sum(x["line_passes"] for x in c) == 1
    assert len({tuple(sorted(x.items())) for x in c}) == len(c)


def test_tune_picks_the_fastest_and_caches(tmp_path):
    path = tmp_path / "t.json"
    reg = dense_region("P", (23, 23, 23), (0.45,) * 3)
    lrs = [FakeLr(reg)] * 3
    def timer(ctx, lrs_, pot, psi, prec, kw):
        o = ctx.opts
        if o["line_passes"]:
            return 9.0
        return (5.0 - (o["zgroup"] == 2) + 0.5 * o["zpot"] - 0.25 * (o["bwd"] == 2)
                - 0.125 * o["tile_order"] + 0.5 * o["persist"])
    ctx = FakeCtx()
    key, opts, times = autotune.tune(ctx, lrs, None, None, bigdft_hip.PREC_BF16X4,
                                     cache=autotune.TuneCache(str(path)), timer=timer)
    assert key == "P:23x23x23:bf16x4"
    assert opts == {"line_passes": 0, "zgroup": 2, "zpot": 0, "bwd": 2, "chunk_mb": 0,
                    "tile_order": 1, "persist": 0}
    assert ctx.opts == opts and len(times) == 14 + len(autotune.CHUNK_MB) + 2
    table = json.loads(path.read_text())
    assert table[key]["options"] == opts and min(table[key]["ms"].values()) == 3.625
    # a second tuner reads the table and times nothing
    ctx2 = FakeCtx()

    def no_timer(*a):
        raise AssertionError("the cache should answer")
    _, opts2, times2 = autotune.tune(ctx2, lrs, None, None, bigdft_hip.PREC_BF16X4,
                                     cache=autotune.TuneCache(str(path)), timer=no_timer)
    assert opts2 == opts and times2 is None and ctx2.opts == opts


def test_bad_cache_file(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text("[1, 2]")
    with pytest.raises(ValueError):
        autotune.TuneCache(str(p))


def test_tune_skips_option_sets_that_do_not_fit(tmp_path):
    reg = dense_region("F", (59, 59, 59), (0.4,) * 3)
    lrs = [FakeLr(reg)]

    def timer(ctx, lrs_, pot, psi, prec, kw):
        if ctx.opts["line_passes"] or ctx.opts["bwd"] == 1:
            raise bigdft_hip.BigDFTError(5, "BDFT_ERR_UNSUPPORTED: LDS")
        return 3.0 + ctx.opts["zgroup"]
    ctx = FakeCtx()
    _, opts, times = autotune.tune(ctx, lrs, None, None, bigdft_hip.PREC_F16X3,
                                   cache=autotune.TuneCache(str(tmp_path / "t.json")),
                                   timer=timer)
    assert opts["bwd"] != 1 and opts["line_passes"] == 0 and opts["zgroup"] == 0
    assert sum(v is None for v in times.values()) == 4

    def other(ctx, *a):
        raise bigdft_hip.BigDFTError(6, "BDFT_ERR_HIP")
    with pytest.raises(bigdft_hip.BigDFTError):
        autotune.tune(FakeCtx(), lrs, None, None, bigdft_hip.PREC_F16X3,
                      cache=autotune.TuneCache(str(tmp_path / "u.json")), timer=other)


def test_stale_cache_entry_is_tuned_again(tmp_path):
    """a cached choice the batch does not fit is dropped and re-tuned"""
    path = tmp_path / "t.json"
    reg = dense_region("F", (40, 40, 40), (0.4,) * 3)
    lrs = [FakeLr(reg)]
    key = autotune.shape_key([reg], bigdft_hip.PREC_F16X3)
    cache = autotune.TuneCache(str(path))
    cache.put(key, {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 1}, {})

    def check(ctx, *a):
        if ctx.opts["bwd"] == 1:
            raise bigdft_hip.BigDFTError(5, "BDFT_ERR_UNSUPPORTED: LDS")

    def timer(ctx, *a):
        if ctx.opts["bwd"] == 1:
            raise bigdft_hip.BigDFTError(5, "BDFT_ERR_UNSUPPORTED: LDS")
        return 1.0 + ctx.opts["zgroup"]
    ctx = FakeCtx()
    _, opts, times = autotune.tune(ctx, lrs, None, None, bigdft_hip.PREC_F16X3,
                                   cache=cache, timer=timer, check=check)
    assert times is not None and opts["bwd"] != 1
    assert json.loads(path.read_text())[key]["options"] == opts


def test_close_candidates_are_timed_again(tmp_path):
    """a candidate that was fast once by chance loses to one that is fast
    every time: the candidates within REFINE of the fastest are timed again
    and the mean decides"""
    reg = dense_region("F", (20, 21, 22), (0.4,) * 3)
    lrs = [FakeLr(reg)] * 2
    calls = {}

    def timer(ctx, lrs_, pot, psi, prec, kw):
        o = ctx.opts
        name = autotune.cand_name(o)
        calls[name] = calls.get(name, 0) + 1
        if o["line_passes"] or o["chunk_mb"] or o["tile_order"] or o["persist"]:
            return 9.0
        if o["zgroup"] == 3 and o["bwd"] == 2:   # lucky first timing
            return 3.50 if calls[name] == 1 else 4.20
        if o["zgroup"] == 2 and o["bwd"] == 2:
            return 3.60
        return 5.0
    ctx = FakeCtx()
    _, opts, times = autotune.tune(ctx, lrs, None, None, bigdft_hip.PREC_F16X3,
                                   cache=autotune.TuneCache(str(tmp_path / "t.json")),
                                   timer=timer)
    assert opts["zgroup"] == 2 and opts["bwd"] == 2 and ctx.opts == opts
    lucky = autotune.cand_name(dict(opts, zgroup=3))
    assert calls[lucky] == 1 + autotune.REFINE_ROUNDS
    assert times[lucky] == pytest.approx((3.5 + autotune.REFINE_ROUNDS * 4.2) /
                                         (1 + autotune.REFINE_ROUNDS))
