"""Host entry points on the GPU against the oracle.

The plain single-series host run (SMA(20), EMA(30), RSI(14) on the seed-0
series of 1,000,000 bars through ct_sma / ct_ema / ct_rsi) under all three
staging modes, with identical NaN ranges; every host single and batch call
of the hot families; host AROON with two periods; the moving-average
selector; the typed (_ex) and _device forms; host candle calls and the
legacy engulfing names; STOCH with long smoothing; the device buffer pool
reusing its buffers."""
import ctypes

import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

F = ctypes.c_void_p


def _p(a):
    return a.ctypes.data


def _scan(g, w, name):
    compare(g, w, tol=1e-5, min_exact=0.999, name=name)


def _exact(g, w, name):
    compare(g, w, tol=0.0, min_exact=1.0, name=name)


def _host1(L, name, x, *params):
    y = np.empty_like(x)
    rc = getattr(L, f"ct_{name}")(_p(x), _p(y), len(x), *params, None)
    assert rc == 0, (name, rc, L.ct_get_last_error())
    return y


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["auto", "pageable", "pinned"])
def test_plain_host_run_1m_bars(dev, mode):
    """ct_sma(20), ct_ema(30), ct_rsi(14) on the seed-0 1,000,000-bar series:
    identical warm-up NaN ranges, SMA bit-identical, EMA/RSI within 1e-5."""
    import tacuda_amd as ta
    from oracle import ORC

    print(f"host run 1M bars mode {mode}", flush=True)
    L = ta.lib()
    assert L.ctx_host_staging(mode) == 0
    try:
        c = ta.synth_ohlcv(1_000_000, seed=0)["close"]
        sma = _host1(L, "sma", c, 20)
        ema = _host1(L, "ema", c, 30)
        rsi = _host1(L, "rsi", c, 14)
        _exact(sma, ORC.sma(c, 20), "host sma20")
        _scan(ema, ORC.ema(c, 30), "host ema30")
        _scan(rsi, ORC.rsi(c, 14), "host rsi14")
        assert np.isnan(sma[:19]).all() and not np.isnan(sma[19:]).any()
        assert np.isnan(ema[:29]).all() and not np.isnan(ema[29:]).any()
        assert np.isnan(rsi[:14]).all() and not np.isnan(rsi[14:]).any()
        # the pool holds the buffers now: repeating the calls allocates nothing
        before = L.ct_device_buffer_pool_allocation_count()
        again = _host1(L, "ema", c, 30)
        _host1(L, "rsi", c, 14)
        assert L.ct_device_buffer_pool_allocation_count() == before
        assert np.array_equal(again.view(np.uint32), ema.view(np.uint32))
    finally:
        L.ctx_host_staging(0)


def test_host_singles_match_oracle(dev):
    import tacuda_amd as ta
    from oracle import ORC

    print("host singles", flush=True)
    L = ta.lib()
    d = ta.synth_ohlcv(6000, seed=3)
    o, h, l, c = d["open"], d["high"], d["low"], d["close"]
    for name, p in (("sma", 20), ("wma", 14), ("sum", 10), ("max", 14), ("min", 14),
                    ("var", 20), ("stddev", 20)):
        _exact(_host1(L, name, c, p), getattr(ORC, name)(c, p), f"host {name}")
    for name, p in (("ema", 20), ("dema", 20), ("tema", 20), ("rsi", 14)):
        _scan(_host1(L, name, c, p), getattr(ORC, name)(c, p), f"host {name}")
    y = np.empty_like(c)
    assert L.ct_t3(_p(c), _p(y), len(c), 5, ctypes.c_float(0.7), None) == 0
    _scan(y, ORC.t3(c, 5, float(np.float32(0.7))), "host t3")
    assert L.ct_kama(_p(c), _p(y), len(c), 30, 2, 30, None) == 0
    _scan(y, ORC.kama(c, 30), "host kama")
    m, s, hh = (np.empty_like(c) for _ in range(3))
    assert L.ct_macd(_p(c), _p(m), _p(s), _p(hh), len(c), 12, 26, 9, 1, None) == 0
    for g, w, k in zip((m, s, hh), ORC.macd(c, 12, 26, 9), ("macd", "signal", "hist")):
        _scan(g, w, f"host {k}")
    u, mid, lo = (np.empty_like(c) for _ in range(3))
    assert L.ct_bbands(_p(c), _p(u), _p(mid), _p(lo), len(c), 20, ctypes.c_float(2.0),
                       ctypes.c_float(1.5), None) == 0
    for g, w, k in zip((u, mid, lo), ORC.bbands(c, 20, 2.0, 1.5), ("up", "mid", "lo")):
        _exact(g, w, f"host bbands {k}")
    for name, p, exact in (("natr", 14, False), ("cci", 14, True), ("adx", 14, False),
                           ("dx", 14, False), ("plus_di", 14, False), ("minus_di", 14, False),
                           ("willr", 14, True)):
        y = np.empty_like(c)
        assert getattr(L, f"ct_{name}")(_p(h), _p(l), _p(c), _p(y), len(c), p, None) == 0
        want = getattr(ORC, name)(h, l, c, p)
        (_exact if exact else _scan)(y, want, f"host {name}")
    assert L.ct_atr(_p(h), _p(l), _p(c), _p(y), len(c), 14, ctypes.c_float(0.0), None) == 0
    _scan(y, ORC.atr(h, l, c, 14), "host atr")
    assert L.ct_trange(_p(h), _p(l), _p(c), _p(y), len(c), None) == 0
    _exact(y, ORC.trange(h, l, c), "host trange")
    k, dd = np.empty_like(c), np.empty_like(c)
    assert L.ct_stochastic(_p(h), _p(l), _p(c), _p(k), _p(dd), len(c), 14, 3, None) == 0
    for g, w, nm in zip((k, dd), ORC.stoch(h, l, c, 14, 3), ("k", "d")):
        _exact(g, w, f"host stoch {nm}")
    for name in ("plus_dm", "minus_dm"):
        assert getattr(L, f"ct_{name}")(_p(h), _p(l), _p(y), len(c), 14, None) == 0
        _scan(y, getattr(ORC, name)(h, l, 14), f"host {name}")
    assert L.ct_aroonosc(_p(h), _p(l), _p(y), len(c), 14, None) == 0
    _exact(y, ORC.aroonosc(h, l, 14), "host aroonosc")


@pytest.mark.parametrize("periods", [(14, 14), (9, 25), (30, 4)])
def test_host_aroon_two_periods(dev, periods):
    import tacuda_amd as ta
    from oracle import ORC

    print(f"host aroon {periods}", flush=True)
    L = ta.lib()
    d = ta.synth_ohlcv(5000, seed=5)
    h, l = d["high"], d["low"]
    up, dn, osc = (np.empty_like(h) for _ in range(3))
    assert L.ct_aroon(_p(h), _p(l), _p(up), _p(dn), _p(osc), len(h), *periods, None) == 0
    for g, w, k in zip((up, dn, osc), ORC.aroon_host(h, l, *periods), ("up", "down", "osc")):
        _exact(g, w, f"aroon {periods} {k}")


def test_host_batch_calls(dev):
    import tacuda_amd as ta
    from oracle import ORC

    print("host batch", flush=True)
    L = ta.lib()
    rows, n = 37, 700
    d = ta.synth_ohlcv(n, rows=rows, seed=9)
    h, l, c = (np.ascontiguousarray(d[k]) for k in ("high", "low", "close"))

    def per_row(fn, *arrs):
        outs = [fn(*[a[r] for a in arrs]) for r in range(rows)]
        if isinstance(outs[0], tuple):
            return tuple(np.stack([o[j] for o in outs]) for j in range(len(outs[0])))
        return np.stack(outs)

    for name, p, exact in (("sma", 20, True), ("ema", 20, False), ("rsi", 14, False),
                           ("wma", 10, True), ("tema", 9, False), ("max", 7, True)):
        y = np.empty_like(c)
        assert getattr(L, f"ct_{name}_batch")(_p(c), _p(y), n, rows, p, 0, None) == 0
        (_exact if exact else _scan)(y, per_row(lambda x: getattr(ORC, name)(x, p), c),
                                     f"batch {name}")
    y = np.empty_like(c)
    assert L.ct_atr_batch(_p(h), _p(l), _p(c), _p(y), n, rows, 14, ctypes.c_float(0), 0,
                          None) == 0
    _scan(y, per_row(lambda a, b, e: ORC.atr(a, b, e, 14), h, l, c), "batch atr")
    assert L.ct_cci_batch(_p(h), _p(l), _p(c), _p(y), n, rows, 14, 0, None) == 0
    _exact(y, per_row(lambda a, b, e: ORC.cci(a, b, e, 14), h, l, c), "batch cci")
    m, s, hh = (np.empty_like(c) for _ in range(3))
    assert L.ct_macd_batch(_p(c), _p(m), _p(s), _p(hh), n, rows, 12, 26, 9, 1, 0, None) == 0
    for g, w, k in zip((m, s, hh), per_row(lambda x: ORC.macd(x, 12, 26, 9), c), "msh"):
        _scan(g, w, f"batch macd {k}")
    u, mid, lo = (np.empty_like(c) for _ in range(3))
    assert L.ct_bbands_batch(_p(c), _p(u), _p(mid), _p(lo), n, rows, 20, ctypes.c_float(2),
                             ctypes.c_float(2), 0, None) == 0
    for g, w, k in zip((u, mid, lo), per_row(lambda x: ORC.bbands(x, 20, 2.0, 2.0), c), "uml"):
        _exact(g, w, f"batch bbands {k}")
    k_, d_ = np.empty_like(c), np.empty_like(c)
    assert L.ct_stochastic_batch(_p(h), _p(l), _p(c), _p(k_), _p(d_), n, rows, 14, 3, 0,
                                 None) == 0
    for g, w, nm in zip((k_, d_), per_row(lambda a, b, e: ORC.stoch(a, b, e, 14, 3), h, l, c),
                        "kd"):
        _exact(g, w, f"batch stoch {nm}")


def test_ma_selector_and_device_aliases(dev):
    """ct_ma_device / ct_ma pick the family the type names (KAMA 2/30, T3
    0.7); the _device forms equal the _d calls bit for bit."""
    import tacuda_amd as ta
    from oracle import ORC

    print("ma selector", flush=True)
    L = ta.lib()
    c = ta.synth_ohlcv(5000, seed=13)["close"]
    x = torch.from_numpy(c).to(dev)
    want = {0: ORC.sma(c, 15), 1: ORC.ema(c, 15), 2: ORC.wma(c, 15), 3: ORC.dema(c, 15),
            4: ORC.tema(c, 15), 6: ORC.kama(c, 15), 8: ORC.t3(c, 15, float(np.float32(0.7)))}
    s = torch.cuda.current_stream().cuda_stream
    for t, w in want.items():
        y = torch.empty_like(x)
        assert L.ct_ma_device(x.data_ptr(), y.data_ptr(), len(c), 15, t, s) == 0
        torch.cuda.synchronize()
        g = y.cpu().numpy()
        (_exact if t in (0, 2) else _scan)(g, w, f"ma type {t}")
        hy = np.empty_like(c)
        assert L.ct_ma(_p(c), _p(hy), len(c), 15, t, None) == 0
        assert np.array_equal(hy.view(np.uint32), g.view(np.uint32)), f"ct_ma type {t}"
    for name in ("sma", "ema", "wma", "dema", "tema", "rsi", "stddev", "var", "sum", "max",
                 "min"):
        a, b = torch.empty_like(x), torch.empty_like(x)
        assert getattr(L, f"ct_{name}_d")(x.data_ptr(), a.data_ptr(), len(c), 14, s) == 0
        assert getattr(L, f"ct_{name}_device")(x.data_ptr(), b.data_ptr(), len(c), 14, s) == 0
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), name


def test_typed_ex_calls(dev):
    """float32 storage with float32 or float64 compute equals the plain
    calls bit for bit, on the device and through the host forms."""
    import tacuda_amd as ta

    print("typed _ex", flush=True)
    L = ta.lib()
    d = ta.synth_ohlcv(4000, seed=17)
    h, l, c = d["high"], d["low"], d["close"]
    s = torch.cuda.current_stream().cuda_stream
    X = {k: torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close")}
    for comp in (1, 2):
        for name in ("sma", "ema", "rsi"):
            a, b = torch.empty_like(X["close"]), torch.empty_like(X["close"])
            assert getattr(L, f"ct_{name}_d")(X["close"].data_ptr(), a.data_ptr(), len(c), 14, s) == 0
            assert getattr(L, f"ct_{name}_ex_d")(X["close"].data_ptr(), b.data_ptr(), len(c), 14,
                                                 1, comp, s) == 0
            torch.cuda.synchronize()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), name
            hy = np.empty_like(c)
            assert getattr(L, f"ct_{name}_ex")(_p(c), _p(hy), len(c), 14, 1, comp, None) == 0
            assert np.array_equal(hy.view(np.uint32), a.cpu().numpy().view(np.uint32)), name
        y0, y1 = np.empty_like(c), np.empty_like(c)
        assert L.ct_atr(_p(h), _p(l), _p(c), _p(y0), len(c), 14, ctypes.c_float(0), None) == 0
        assert L.ct_atr_ex(_p(h), _p(l), _p(c), _p(y1), len(c), 14, ctypes.c_float(0), 1, comp,
                           None) == 0
        assert np.array_equal(y0.view(np.uint32), y1.view(np.uint32))
        m0 = [np.empty_like(c) for _ in range(3)]
        m1 = [np.empty_like(c) for _ in range(3)]
        assert L.ct_macd(_p(c), *map(_p, m0), len(c), 12, 26, 9, 1, None) == 0
        assert L.ct_macd_ex(_p(c), *map(_p, m1), len(c), 12, 26, 9, 1, 1, comp, None) == 0
        b0 = [np.empty_like(c) for _ in range(3)]
        b1 = [np.empty_like(c) for _ in range(3)]
        f2 = ctypes.c_float(2.0)
        assert L.ct_bbands(_p(c), *map(_p, b0), len(c), 20, f2, f2, None) == 0
        assert L.ct_bbands_ex(_p(c), *map(_p, b1), len(c), 20, f2, f2, 1, comp, None) == 0
        for u, v in zip(m0 + b0, m1 + b1):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    # element sizes of the typed copies
    for t, dt in ((0, np.float16), (1, np.float32), (2, np.float64)):
        src = np.arange(1000, dtype=dt)
        back = np.zeros_like(src)
        buf = torch.empty(1000 * src.itemsize, dtype=torch.uint8, device=dev)
        assert L.ct_device_upload_ex(buf.data_ptr(), _p(src), 1000, t, None) == 0
        assert L.ct_device_download_ex(_p(back), buf.data_ptr(), 1000, t, None) == 0
        assert np.array_equal(src, back)


def test_host_candle_calls(dev):
    import tacuda_amd as ta
    from oracle import ORC
    from tests.candle_data import quantized

    print("host candles", flush=True)
    L = ta.lib()
    o, h, l, c = quantized(3000, seed=2)
    rows, n = 4, 750
    for name in ta.ops.cdl_names():
        y = np.empty_like(c)
        assert getattr(L, f"ct_cdl_{name}")(_p(o), _p(h), _p(l), _p(c), _p(y), len(c), None) == 0
        want = ORC.candle(name, o, h, l, c).astype(np.float32)
        assert np.array_equal(y, want), name
        yb = np.empty_like(c)
        assert getattr(L, f"ct_cdl_{name}_batch")(_p(o), _p(h), _p(l), _p(c), _p(yb), n, rows, 0,
                                                   None) == 0
        wb = np.concatenate([ORC.candle(name, *(a[r * n:(r + 1) * n] for a in (o, h, l, c)))
                             for r in range(rows)]).astype(np.float32)
        assert np.array_equal(yb, wb), name + " batch"


def test_legacy_engulfing_names(dev):
    import tacuda_amd as ta
    from oracle import ORC
    from tests.candle_data import quantized

    print("legacy engulfing", flush=True)
    L = ta.lib()
    o, h, l, c = quantized(4000, seed=4)
    rows, n = 5, 800
    want = ORC.engulfing_legacy(o, h, l, c)
    wantb = np.concatenate([ORC.engulfing_legacy(*(a[r * n:(r + 1) * n] for a in (o, h, l, c)))
                            for r in range(rows)])
    assert (want != 0).sum() > 50
    T = [torch.from_numpy(a).to(dev) for a in (o, h, l, c)]
    s = torch.cuda.current_stream().cuda_stream
    for name in ("bullish_engulfing", "bearish_engulfing"):
        y = np.empty_like(c)
        assert getattr(L, f"ct_cdl_{name}")(*map(_p, (o, h, l, c, y)), len(c), None) == 0
        assert np.array_equal(y, want), name
        assert getattr(L, f"ct_cdl_{name}_batch")(*map(_p, (o, h, l, c, y)), n, rows, 0, None) == 0
        assert np.array_equal(y, wantb), name + " batch"
        yd = torch.empty_like(T[3])
        assert getattr(L, f"ct_cdl_{name}_d")(*[t.data_ptr() for t in T], yd.data_ptr(), len(c),
                                              s) == 0
        torch.cuda.synchronize()
        assert np.array_equal(yd.cpu().numpy(), want), name + " _d"
        assert getattr(L, f"ct_cdl_{name}_batch_d")(*[t.data_ptr() for t in T], yd.data_ptr(), n,
                                                    rows, 0, s) == 0
        torch.cuda.synchronize()
        assert np.array_equal(yd.cpu().numpy(), wantb), name + " batch_d"


@pytest.mark.parametrize("kd", [(14, 10), (5, 30), (30, 12)])
def test_stoch_long_smoothing(dev, kd):
    """STOCH with dPeriod above 9 (the reference accepts any dPeriod > 0):
    single series, dense batch and ragged layout bit-identical to the
    oracle."""
    import tacuda_amd as ta
    from oracle import ORC
    from tests.test_gpu_ragged import _cat_want, _layout

    kp, dp = kd
    print(f"stoch {kp},{dp}", flush=True)
    d = ta.synth_ohlcv(20000, seed=21)
    T = [torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close")]
    k, dd = ta.ops.stochastic(*T, kp, dp)
    for g, w, nm in zip((k, dd), ORC.stoch(d["high"], d["low"], d["close"], kp, dp), "kd"):
        _exact(g.cpu().numpy(), w, f"stoch {kd} {nm}")
    b = ta.synth_ohlcv(500, rows=30, seed=22)
    B = [torch.from_numpy(np.ascontiguousarray(b[k])).to(dev) for k in ("high", "low", "close")]
    k, dd = ta.ops.stochastic(*B, kp, dp)
    for r in (0, 13, 29):
        for g, w, nm in zip((k[r], dd[r]), ORC.stoch(b["high"][r], b["low"][r], b["close"][r],
                                                     kp, dp), "kd"):
            _exact(g.cpu().numpy(), w, f"stoch batch {kd} row {r} {nm}")
    lengths = [40, 300, 1, 2500, 77]
    off, cat = _layout(lengths, seed=23)
    lay = ta.ops.Ragged(off)
    R = [torch.from_numpy(cat[k]).to(dev) for k in ("high", "low", "close")]
    k, dd = lay.stochastic(*R, kp, dp)
    want = _cat_want(off, lambda a, bb, cc: ORC.stoch(a, bb, cc, kp, dp), cat["high"], cat["low"],
                     cat["close"])
    _exact(k.cpu().numpy(), want[0], f"stoch ragged {kd} k")
    _exact(dd.cpu().numpy(), want[1], f"stoch ragged {kd} d")
    lay.close()


def test_ma_d_and_batch_forms(dev):
    """ct_ma_d, ct_ma_batch_d and ct_ma_batch pick the same family by type."""
    import tacuda_amd as ta
    from oracle import ORC

    print("ma batch forms", flush=True)
    L = ta.lib()
    rows, n = 11, 600
    b = ta.synth_ohlcv(n, rows=rows, seed=41)["close"]
    b = np.ascontiguousarray(b)
    B = torch.from_numpy(b).to(dev)
    s = torch.cuda.current_stream().cuda_stream
    for t, fn in ((0, lambda x: ORC.sma(x, 9)), (3, lambda x: ORC.dema(x, 9)),
                  (5, lambda x: ORC.trima(x, 9)), (6, lambda x: ORC.kama(x, 9))):
        y = torch.empty_like(B)
        assert L.ct_ma_batch_d(B.data_ptr(), y.data_ptr(), n, rows, 9, t, 0, s) == 0
        torch.cuda.synchronize()
        g = y.cpu().numpy()
        for r in range(rows):
            (_exact if t in (0, 5) else _scan)(g[r], fn(b[r]), f"ma batch type {t} row {r}")
        hy = np.empty_like(b)
        assert L.ct_ma_batch(_p(b), _p(hy), n, rows, 9, t, 0, None) == 0
        assert np.array_equal(hy.view(np.uint32), g.view(np.uint32))
        y1 = torch.empty_like(B[0])
        assert L.ct_ma_d(B[0].data_ptr(), y1.data_ptr(), n, 9, t, s) == 0
        torch.cuda.synchronize()
        assert np.array_equal(y1.cpu().numpy().view(np.uint32), g[0].view(np.uint32))
