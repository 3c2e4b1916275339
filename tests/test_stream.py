"""Streaming API on the host: appending bars in arbitrary chunks reproduces
the oracle's full-series result bit for bit (warm-up NaNs included), for
every streamed family; handle lifecycle and batch streams on the CPU."""
import numpy as np
import pytest

import tacuda_amd as ta
from oracle import ORC
from tacuda_amd.stream import Stream, StreamBatch

F32 = lambda v: float(np.float32(v))  # noqa: E731

CASES = {
    "sma": ((20,), lambda h, l, c, v: ORC.sma(c, 20)),
    "sma1": ((1,), lambda h, l, c, v: ORC.sma(c, 1)),
    "ema": ((20,), lambda h, l, c, v: ORC.ema(c, 20)),
    "ema1": ((1,), lambda h, l, c, v: ORC.ema(c, 1)),
    "wma": ((14,), lambda h, l, c, v: ORC.wma(c, 14)),
    "rsi": ((14,), lambda h, l, c, v: ORC.rsi(c, 14)),
    "rsi1": ((1,), lambda h, l, c, v: ORC.rsi(c, 1)),
    "atr": ((14,), lambda h, l, c, v: ORC.atr(h, l, c, 14)),
    "atr1": ((1,), lambda h, l, c, v: ORC.atr(h, l, c, 1)),
    "natr": ((14,), lambda h, l, c, v: ORC.natr(h, l, c, 14)),
    "obv": ((), lambda h, l, c, v: ORC.obv(c, v)),
    "ad": ((), lambda h, l, c, v: ORC.ad(h, l, c, v)),
    "macd": ((12, 26, 9), lambda h, l, c, v: ORC.macd(c, 12, 26, 9)),
    "macd_rev_sig1": ((26, 12, 1), lambda h, l, c, v: ORC.macd(c, 26, 12, 1)),
    "bbands": ((20, 2.0, 2.0), lambda h, l, c, v: ORC.bbands(c, 20, 2.0, 2.0)),
    "bbands_asym": ((20, 2.0, 1.5), lambda h, l, c, v: ORC.bbands(c, 20, 2.0, F32(1.5))),
}


def _family(name):
    return name.split("1")[0].split("_")[0] if name not in ("sma1", "ema1", "rsi1", "atr1") \
        else name[:-1]


def _data(n, seed=0, prefix=0, hole=None):
    d = ta.synth_ohlcv(n, seed=seed)
    d = {k: v.copy() for k, v in d.items()}
    for k in ("open", "high", "low", "close"):
        d[k][:prefix] = np.nan
        if hole is not None:
            d[k][hole] = np.nan
    return d


def _bits(a, b, name):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{name}: NaN masks differ"
    ok = ~np.isnan(a)
    bad = np.flatnonzero(a[ok].view(np.uint32) != b[ok].view(np.uint32))
    assert bad.size == 0, f"{name}: {bad.size} values differ, first {a[ok][bad[:3]]} vs {b[ok][bad[:3]]}"


def _stream_all(name, d, chunks):
    prm, _ = CASES[name]
    s = Stream(_family(name), *prm)
    outs = []
    i = 0
    for ch in chunks:
        sl = slice(i, i + ch)
        outs.append(s.append(high=d["high"][sl], low=d["low"][sl], close=d["close"][sl],
                             volume=d["volume"][sl]))
        i += ch
    s.close()
    if isinstance(outs[0], tuple):
        return tuple(np.concatenate([o[j] for o in outs]) for j in range(len(outs[0])))
    return np.concatenate(outs)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("variant", ["plain", "prefix", "hole"])
def test_stream_matches_full_series(name, variant):
    n = 3000
    kw = {"plain": {}, "prefix": {"prefix": 33}, "hole": {"hole": 1700}}[variant]
    d = _data(n, seed=5, **kw)
    rng = np.random.default_rng(len(name))
    chunks = []
    left = n
    while left:
        ch = int(min(left, rng.choice([1, 1, 2, 7, 50, 333])))
        chunks.append(ch)
        left -= ch
    got = _stream_all(name, d, chunks)
    want = CASES[name][1](d["high"], d["low"], d["close"], d["volume"])
    got = got if isinstance(got, tuple) else (got,)
    want = want if isinstance(want, tuple) else (want,)
    for j, (g, w) in enumerate(zip(got, want)):
        _bits(g, w, f"{name}[{j}] {variant}")


def test_stream_lifecycle_and_queries():
    s = Stream("macd", 12, 26, 9)
    assert s.lookback == 33 and s.outputs == 3 and s.bars == 0
    s.append(close=np.arange(40, dtype=np.float32))
    assert s.bars == 40
    s.reset()
    assert s.bars == 0
    assert Stream("rsi", 14).lookback == 14
    assert Stream("obv").lookback == 0
    L = ta.lib()
    h = s._h
    s.close()
    out = np.zeros(1, np.float32)
    x = np.ones(1, np.float32)
    # a destroyed token is rejected, never dereferenced
    assert L.ct_ema_stream_append(h, x.ctypes.data, 1, out.ctypes.data) == 4
    assert L.ct_stream_destroy(h) == 4
    assert L.ct_stream_destroy(None) == 0
    e = Stream("ema", 5)
    # appending through another family's entry point is rejected
    assert L.ct_sma_stream_append(e._h, x.ctypes.data, 1, out.ctypes.data) == 4
    with pytest.raises(ta.TacudaError):
        Stream("ema", 0)
    with pytest.raises(ta.TacudaError):
        StreamBatch("ema", 0, 5)


@pytest.mark.parametrize("family,prm", [("ema", (20,)), ("rsi", (14,)), ("atr", (14,)),
                                        ("bbands", (20, 2.0, 2.0)), ("macd", (12, 26, 9)),
                                        ("ad", ()), ("wma", (10,))])
def test_batch_cpu_equals_single_streams(family, prm):
    rows, n = 37, 400
    d = ta.synth_ohlcv(n, rows=rows, seed=11)
    b = StreamBatch(family, rows, *prm, exec="cpu")
    outs = []
    for lo, hi in ((0, 1), (1, 150), (150, 151), (151, 400)):
        sl = (slice(None), slice(lo, hi))
        outs.append(b.append(high=d["high"][sl], low=d["low"][sl], close=d["close"][sl],
                             volume=d["volume"][sl]))
    assert b.bars == n and b.series == rows
    for r in (0, 5, 36):
        s = Stream(family, *prm)
        want = s.append(high=d["high"][r], low=d["low"][r], close=d["close"][r],
                        volume=d["volume"][r])
        want = want if isinstance(want, tuple) else (want,)
        for j in range(len(want)):
            got = np.concatenate([(o if isinstance(o, tuple) else (o,))[j][r] for o in outs])
            _bits(got, want[j], f"{family} row {r} out {j}")
