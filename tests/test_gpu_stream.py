"""Batch streams on the GPU against the oracle: every append (host arrays
and device pointers) equals the oracle's full-series result over the
concatenated history, bit for bit, for all ten streamed families; the
HBM-resident state machines also equal the host ones (including switching
execution mid-stream), and the 65,536-symbol EMA(20)/RSI(14)/ATR(14) tick
matches the oracle on every symbol."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FAMS = [("ema", (20,)), ("rsi", (14,)), ("atr", (14,)), ("natr", (14,)), ("sma", (20,)),
        ("wma", (10,)), ("bbands", (20, 2.0, 1.5)), ("macd", (12, 26, 9)), ("obv", ()),
        ("ad", ())]


def _tup(x):
    return x if isinstance(x, tuple) else (x,)


def _oracle(family, prm, h, l, c, v):
    """The oracle's full-series result of a streamed family (the stream
    contract, include/tacuda_stream.h:8-20: appending bars in any chunks
    gives the full-series values)."""
    from oracle import ORC

    f32 = lambda x: float(np.float32(x))  # noqa: E731
    if family in ("ema", "rsi", "sma", "wma"):
        return _tup(getattr(ORC, family)(c, *prm))
    if family in ("atr", "natr"):
        return _tup(getattr(ORC, family)(h, l, c, *prm))
    if family == "bbands":
        p, up, dn = prm
        return ORC.bbands(c, p, f32(up), f32(dn))
    if family == "macd":
        return ORC.macd(c, *prm)
    if family == "obv":
        return _tup(ORC.obv(c, v))
    return _tup(ORC.ad(h, l, c, v))


def _eq(a, b, name):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), name
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), name


@pytest.mark.parametrize("family,prm", FAMS, ids=[f[0] for f in FAMS])
def test_gpu_batch_equals_cpu(dev, family, prm):
    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    print(f"stream batch {family}", flush=True)
    rows, n = 3000, 120
    d = ta.synth_ohlcv(n, rows=rows, seed=17)
    d["close"][7, 40] = np.nan
    cpu = StreamBatch(family, rows, *prm, exec="cpu")
    gpu = StreamBatch(family, rows, *prm, exec="gpu")
    got = []
    for lo, hi in ((0, 1), (1, 60), (60, 61), (61, 120)):
        sl = (slice(None), slice(lo, hi))
        kw = dict(high=d["high"][sl], low=d["low"][sl], close=d["close"][sl],
                  volume=d["volume"][sl])
        g = _tup(gpu.append(**kw))
        got.append(g)
        for j, (a, b) in enumerate(zip(_tup(cpu.append(**kw)), g)):
            _eq(b, a, f"{family} [{lo},{hi}) out {j}")
    assert gpu.bars == n
    full = [np.concatenate([g[j] for g in got], axis=1) for j in range(len(got[0]))]
    for r in (0, 7, 1500, 2999):
        want = _oracle(family, prm, d["high"][r], d["low"][r], d["close"][r], d["volume"][r])
        for j, w in enumerate(want):
            _eq(full[j][r], w, f"{family} row {r} out {j} vs oracle")


@pytest.mark.parametrize("family,prm", FAMS, ids=[f[0] for f in FAMS])
def test_gpu_device_append_matches_oracle(dev, family, prm):
    """ctx_stream_batch_append_d: device tensors in and out, states in HBM;
    every append's outputs equal the oracle's full-series values."""
    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    print(f"stream append_d {family}", flush=True)
    rows, n = 1024, 90
    d = ta.synth_ohlcv(n, rows=rows, seed=23)
    d["close"][3, 50] = np.nan
    D = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev)
         for k in ("high", "low", "close", "volume")}
    b = StreamBatch(family, rows, *prm, exec="gpu")
    got = []
    for lo, hi in ((0, 30), (30, 31), (31, 32), (32, 90)):
        kw = {k: D[k][:, lo:hi].contiguous() for k in D}
        outs = _tup(b.append_d(high=kw["high"], low=kw["low"], close=kw["close"],
                               volume=kw["volume"]))
        got.append([o.reshape(rows, hi - lo).cpu().numpy() for o in outs])
    full = [np.concatenate([g[j] for g in got], axis=1) for j in range(len(got[0]))]
    for r in (0, 3, 511, 1023):
        want = _oracle(family, prm, d["high"][r], d["low"][r], d["close"][r], d["volume"][r])
        for j, w in enumerate(want):
            _eq(full[j][r], w, f"{family} append_d row {r} out {j}")
    b.close()


def test_tick_65536_symbols_matches_oracle(dev):
    """The streaming tick at its real shape: 65,536 symbols' EMA(20),
    RSI(14) and ATR(14) states after a 40-bar history, advanced one bar per
    tick (three ticks from host arrays, three from device tensors); every
    symbol's outputs equal the oracle over its concatenated history."""
    import tacuda_amd as ta
    from oracle import ORC
    from tacuda_amd.stream import StreamBatch

    print("tick 65536 symbols", flush=True)
    S, hist, ticks = 65536, 40, 6
    n = hist + ticks
    d = ta.synth_ohlcv(n, rows=S, seed=29)
    fams = [("ema", (20,)), ("rsi", (14,)), ("atr", (14,))]
    bs = [StreamBatch(f, S, *p, exec="gpu") for f, p in fams]
    got = [[] for _ in bs]

    def cols(lo, hi):
        return {k: np.ascontiguousarray(d[k][:, lo:hi]) for k in ("high", "low", "close")}

    for b, g in zip(bs, got):
        g.append(b.append(**cols(0, hist)))
    for t in range(hist, hist + 3):
        for b, g in zip(bs, got):
            g.append(b.append(**cols(t, t + 1)))
    D = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")}
    for t in range(hist + 3, n):
        for b, g in zip(bs, got):
            kw = {k: D[k][:, t].contiguous() for k in D}
            g.append(b.append_d(**kw).cpu().numpy().reshape(S, 1))
    for b in bs:
        assert b.bars == n
        b.close()
    for (fam, prm), g in zip(fams, got):
        full = np.concatenate(g, axis=1)
        for r in range(S):
            if fam == "atr":
                w = ORC.atr(d["high"][r], d["low"][r], d["close"][r], *prm)
            else:
                w = getattr(ORC, fam)(d["close"][r], *prm)
            if not np.array_equal(full[r].view(np.uint32), w.view(np.uint32)):
                _eq(full[r], w, f"tick {fam} symbol {r}")


def test_device_append_and_exec_switch(dev):
    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    print("stream device append", flush=True)
    rows, n = 65536, 6
    d = ta.synth_ohlcv(n, rows=rows, seed=19)
    ref = StreamBatch("rsi", rows, 3, exec="cpu")
    want = ref.append(close=d["close"])
    b = StreamBatch("rsi", rows, 3, exec="auto")
    c = torch.from_numpy(np.ascontiguousarray(d["close"])).to(dev)
    got = []
    # bar 0 on the host (AUTO picks the GPU at this size), bars 1-3 on the
    # device, bar 4 forced to the CPU, bar 5 back on the GPU
    got.append(b.append(close=d["close"][:, 0:1]))
    for i in (1, 2, 3):
        col = c[:, i].contiguous()
        got.append(b.append_d(close=col).cpu().numpy()[:, None])
    b.set_exec("cpu")
    got.append(b.append(close=d["close"][:, 4:5]))
    b.set_exec("gpu")
    got.append(b.append(close=d["close"][:, 5:6]))
    full = np.concatenate(got, axis=1)
    _eq(full, want, "rsi mixed execution")
    assert b.bars == n
    from oracle import ORC

    for r in (0, 1, 32768, 65535):
        _eq(full[r], ORC.rsi(d["close"][r], 3), f"rsi mixed execution row {r} vs oracle")
