"""Pins of the CPU oracle (oracle/oracle.c), runnable without a GPU.

Each family of the oracle is checked by something other than a retyped
formula: the reference's own TA-Lib 0.8.3 host build (oracle/_ref), the
reference's TA-Lib golden fixture, the reference's known-answer tests, closed
forms, invariants, and structurally different numpy restatements.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import ORC, REF, ref

HERE = os.path.dirname(os.path.abspath(__file__))


def _walk(n, seed, drift=0.0):
    rng = np.random.default_rng(seed)
    c = 100 * np.exp(np.cumsum(rng.normal(drift, 0.01, n)))
    o = np.concatenate([[100.0], c[:-1]])
    h = np.maximum(o, c) * (1 + np.abs(rng.normal(0, 0.005, n)))
    l = np.minimum(o, c) * (1 - np.abs(rng.normal(0, 0.005, n)))
    return [a.astype(np.float32) for a in (o, h, l, c)]


DATASETS = {
    "rw0": _walk(3000, 0),
    "rw1": _walk(3000, 1, 0.001),
    "short": _walk(40, 5),
    "flat": [np.full(500, 123.45, np.float32)] * 4,
    "zeros": [np.sin(np.arange(800) * 0.1).astype(np.float32) * 50] * 4,
}


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN masks differ"
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), (
        f"max diff {np.max(np.abs(a[ok].astype(np.float64) - b[ok]))}")


needs_ref = pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (no reference checkout)")

VENDORED = [
    ("sma", lambda F, o, h, l, c: F.sma(c, 20)),
    ("sma1", lambda F, o, h, l, c: F.sma(c, 1)),
    ("ema", lambda F, o, h, l, c: F.ema(c, 20)),
    ("ema2", lambda F, o, h, l, c: F.ema(c, 2)),
    ("wma", lambda F, o, h, l, c: F.wma(c, 20)),
    ("dema", lambda F, o, h, l, c: F.dema(c, 20)),
    ("tema", lambda F, o, h, l, c: F.tema(c, 20)),
    ("t3", lambda F, o, h, l, c: F.t3(c, 5, 0.7)),
    ("kama", lambda F, o, h, l, c: F.kama(c, 30)),
    ("rsi", lambda F, o, h, l, c: F.rsi(c, 14)),
    ("rma", lambda F, o, h, l, c: F.rma(c, 14)),
    ("atr", lambda F, o, h, l, c: F.atr(h, l, c, 14)),
    ("atr1", lambda F, o, h, l, c: F.atr(h, l, c, 1)),
    ("trange", lambda F, o, h, l, c: F.trange(h, l, c)),
    ("macd", lambda F, o, h, l, c: F.macd(c, 12, 26, 9)),
    ("macd_rev", lambda F, o, h, l, c: F.macd(c, 26, 12, 9)),
    ("macd_sig1", lambda F, o, h, l, c: F.macd(c, 5, 9, 1)),
    ("bbands", lambda F, o, h, l, c: F.bbands(c, 20, 2.0, 2.0)),
    ("bbands_asym", lambda F, o, h, l, c: F.bbands(c, 20, 2.0, 1.5)),
    ("var", lambda F, o, h, l, c: F.var(c, 20)),
    ("stddev", lambda F, o, h, l, c: F.stddev(c, 20)),
    ("stoch", lambda F, o, h, l, c: F.stoch(h, l, c, 14, 3)),
    ("stoch5_3", lambda F, o, h, l, c: F.stoch(h, l, c, 5, 3)),
]


@needs_ref
@pytest.mark.parametrize("ds", sorted(DATASETS))
@pytest.mark.parametrize("name,fn", VENDORED, ids=[v[0] for v in VENDORED])
def test_oracle_matches_reference_talib(ds, name, fn):
    """Bit equality (float32) with the reference's TA-Lib 0.8.3 build."""
    o, h, l, c = DATASETS[ds]
    got, want = fn(ORC, o, h, l, c), fn(REF, o, h, l, c)
    got = got if isinstance(got, tuple) else (got,)
    want = want if isinstance(want, tuple) else (want,)
    for g, w in zip(got, want):
        if name.startswith("stoch"):
            # TA_STOCH smooths with TA_SMA's add-then-subtract running total over
            # non-exact doubles; its residue (~1e-14 where a window of %K is all 0)
            # differs from the direct window sum the reference semantics use.
            assert np.array_equal(np.isnan(g), np.isnan(w))
            ok = ~np.isnan(w)
            np.testing.assert_allclose(g[ok], w[ok], rtol=0, atol=1e-9)
        else:
            _bits_equal(g, w)


def test_rma_golden_fixture():
    """TA-Lib 0.8.3 RMA(14) values from the reference's fixture file."""
    d = json.load(open(os.path.join(HERE, "golden", "talib083_rma14.json")))
    c = np.array(d["close"], np.float32)
    want = np.array([np.nan if v is None else v for v in d["rma14"]])
    got = ORC.rma(c, 14)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    # TA-Lib computed in double; the oracle's float32 output is its rounding
    assert np.array_equal(got[ok], want[ok].astype(np.float32))


def test_rma_equals_atr_of_trange():
    """TA_RMA(TA_TRANGE) == TA_ATR bit for bit (ta_ATR.c:103-110)."""
    o, h, l, c = DATASETS["rw0"]
    tr = ORC.trange(h, l, c)
    rma = ORC.rma(tr[1:], 14)
    atr = ORC.atr(h, l, c, 14)
    # ATR's seed spans TR[1..p]; RMA over TR[1:] seeds over the same window.
    # The RMA input is float32-rounded TR, so compare to the ATR of the same.
    assert np.allclose(rma, atr[1:], rtol=1e-6, equal_nan=True)


@pytest.mark.parametrize("direction,expected", [(1.0, 200.0 / 3.0), (-1.0, 100.0 / 3.0)])
def test_rsi_pivot_known_answer(direction, expected):
    """Reference tests/cpp/test_rsi.cpp:26-53."""
    n, p = 96, 7
    x = np.full(n, 100.0 + direction, np.float32)
    x[0] = 100.0
    x[1] = 100.0 + 2.0 * direction
    got = ORC.rsi(x, p)
    assert np.isnan(got[:p]).all()
    assert (got[p:] == np.float32(expected)).all()


@pytest.mark.parametrize("k,d", [(1, 1), (5, 3), (65, 3), (14, 9)])
def test_stoch_ramp_known_answer(k, d):
    """Reference tests/cpp/test_indicator_audit_regressions.cpp:64-100."""
    minimum = k + 2 * (d - 1)
    n = minimum + 3
    c = np.arange(1, n + 1, dtype=np.float32)
    kk, dd = ORC.stoch(c + 1, c - 1, c, k, d)
    assert np.isnan(kk[: minimum - 1]).all() and np.isnan(dd[: minimum - 1]).all()
    np.testing.assert_allclose(kk[minimum - 1:], 100.0 * k / (k + 1), rtol=1e-6)
    np.testing.assert_allclose(dd[minimum - 1:], 100.0 * k / (k + 1), rtol=1e-6)


def test_flat_series_degenerate_values():
    """docs/IMPLEMENTATION_NOTES.md:46-58 degenerate-input conventions."""
    x = np.full(300, 123.45, np.float32)
    assert (ORC.rsi(x, 14)[14:] == 50.0).all()
    k, d = ORC.stoch(x, x, x, 14, 3)
    assert (k[17:] == 0).all() and (d[17:] == 0).all()
    assert (ORC.willr(x, x, x, 14)[13:] == 0).all()
    assert (ORC.cci(x, x, x, 14)[13:] == 0).all()
    assert (ORC.stddev(x, 20)[19:] == 0).all()
    assert (ORC.var(x, 20)[19:] == 0).all()
    assert (ORC.adx(x, x, x, 14)[27:] == 0).all()
    assert (ORC.atr(x, x, x, 14)[14:] == 0).all()
    u, m, l = ORC.bbands(x, 20, 2.0, 2.0)
    assert (u[19:] == x[0]).all() and (l[19:] == x[0]).all()
    assert (ORC.ema(x, 20)[19:] == x[0]).all()
    assert (ORC.kama(x, 10)[10:] == x[0]).all()


def test_ema_closed_form_step():
    """EMA of a unit step after an SMA seed: y_t = 1 - beta^(t-s)*(1 - seed)."""
    p, n = 9, 200
    x = np.zeros(n, np.float32)
    x[20:] = 1.0
    got = ORC.ema(x, p).astype(np.float64)
    beta = (p - 1) / (p + 1)
    t = np.arange(n)
    want = np.where(t < 20, 0.0, 1.0 - beta ** (t - 19))
    np.testing.assert_allclose(got[p - 1:], want[p - 1:], atol=1e-7)


def test_sma_wma_exact_rationals():
    """Window means against exact rational arithmetic."""
    o, h, l, c = DATASETS["short"]
    p = 7
    sma, wma = ORC.sma(c, p), ORC.wma(c, p)
    for t in range(p - 1, len(c)):
        win = [Fraction(float(v)) for v in c[t - p + 1: t + 1]]
        assert sma[t] == np.float32(float(sum(win) / p))
        wsum = sum(Fraction(i + 1) * v for i, v in enumerate(win))
        assert wma[t] == np.float32(float(wsum / Fraction(p * (p + 1), 2)))


def test_var_two_pass_bruteforce():
    o, h, l, c = DATASETS["rw1"]
    p = 20
    got = ORC.var(c, p).astype(np.float64)
    w = np.lib.stride_tricks.sliding_window_view(c.astype(np.float64), p)
    want = w.var(axis=1)
    np.testing.assert_allclose(got[p - 1:], want, rtol=2e-6, atol=1e-9)


# ------------------------------------------------ families not vendored ---

def _np_dm(h, l):
    h = h.astype(np.float64)
    l = l.astype(np.float64)
    up = np.diff(h)
    dn = -np.diff(l)
    pdm = np.where((up > 0) & (up > dn), up, 0.0)
    mdm = np.where((dn > 0) & (dn > up), dn, 0.0)
    return pdm, mdm


def _np_tr(h, l, c):
    h, l, c = (a.astype(np.float64) for a in (h, l, c))
    return np.maximum.reduce([h[1:] - l[1:], np.abs(c[:-1] - h[1:]), np.abs(c[:-1] - l[1:])])


def test_dm_di_dx_adx_restated():
    """Directional movement from a vectorized Wilder restatement."""
    o, h, l, c = DATASETS["rw0"]
    p = 14
    pdm, mdm = _np_dm(h, l)
    tr = _np_tr(h, l, c)

    def wilder(u):  # u indexed by bar 1..n-1
        out = np.full(len(c), np.nan)
        s = u[: p - 1].sum()
        out[p - 1] = s
        for t in range(p, len(c)):
            s = s - s / p + u[t - 1]
            out[t] = s
        return out

    sp, sm, st = wilder(pdm), wilder(mdm), wilder(tr)
    np.testing.assert_allclose(ORC.plus_dm(h, l, p)[p - 1:], sp[p - 1:], rtol=1e-6)
    np.testing.assert_allclose(ORC.minus_dm(h, l, p)[p - 1:], sm[p - 1:], rtol=1e-6)
    pdi, mdi = 100 * sp / st, 100 * sm / st
    np.testing.assert_allclose(ORC.plus_di(h, l, c, p)[p:], pdi[p:], rtol=1e-5)
    np.testing.assert_allclose(ORC.minus_di(h, l, c, p)[p:], mdi[p:], rtol=1e-5)
    dx = 100 * np.abs(pdi - mdi) / (pdi + mdi)
    np.testing.assert_allclose(ORC.dx(h, l, c, p)[p:], dx[p:], rtol=1e-5, atol=1e-5)
    adx = np.full(len(c), np.nan)
    adx[2 * p - 1] = dx[p: 2 * p].mean()
    for t in range(2 * p, len(c)):
        adx[t] = (adx[t - 1] * (p - 1) + dx[t]) / p
    got = ORC.adx(h, l, c, p)
    assert np.isnan(got[: 2 * p - 1]).all()
    np.testing.assert_allclose(got[2 * p - 1:], adx[2 * p - 1:], rtol=1e-5, atol=1e-5)


def test_dx_scale_invariant_and_ranges():
    o, h, l, c = DATASETS["rw1"]
    a = ORC.dx(h, l, c, 14)
    b = ORC.dx(h * 4, l * 4, c * 4, 14)  # power-of-two scaling is exact
    np.testing.assert_array_equal(a, b)
    adx = ORC.adx(h, l, c, 14)
    ok = ~np.isnan(adx)
    assert (adx[ok] >= 0).all() and (adx[ok] <= 100).all()


def test_willr_cci_aroon_restated():
    o, h, l, c = DATASETS["rw0"]
    p = 14
    H = np.lib.stride_tricks.sliding_window_view(h, p).max(axis=1).astype(np.float64)
    L = np.lib.stride_tricks.sliding_window_view(l, p).min(axis=1).astype(np.float64)
    wr = np.clip((H - c[p - 1:]) / (H - L) * -100.0, -100, 0)
    np.testing.assert_allclose(ORC.willr(h, l, c, p)[p - 1:], wr, rtol=1e-6)
    tp = (h.astype(np.float64) + l + c) / 3.0
    w = np.lib.stride_tricks.sliding_window_view(tp, p)
    sma = w.mean(axis=1)
    md = np.abs(w - sma[:, None]).mean(axis=1)
    cci = (tp[p - 1:] - sma) / (0.015 * md)
    np.testing.assert_allclose(ORC.cci(h, l, c, p)[p - 1:], cci, rtol=1e-5, atol=1e-4)
    dn, up = ORC.aroon(h, l, p)
    for t in range(p, len(c)):
        wh = h[t - p: t + 1][::-1]  # newest first: argmax picks the newest tie
        wl = l[t - p: t + 1][::-1]
        assert up[t] == np.float32(100.0 / p * (p - int(np.argmax(wh))))
        assert dn[t] == np.float32(100.0 / p * (p - int(np.argmin(wl))))
    osc = ORC.aroonosc(h, l, p)
    np.testing.assert_allclose(osc[p:], up[p:] - dn[p:], atol=1e-4)


def test_min_max_nan_rules():
    x = np.arange(30, dtype=np.float32)
    x[:4] = np.nan
    x[10] = np.nan
    mx, mn = ORC.max(x, 5), ORC.min(x, 5)
    for t in range(4, 30):
        s = t - 4
        win = x[s: t + 1]
        if s < 4:
            assert np.isnan(mx[t]) and np.isnan(mn[t])
        else:
            assert mx[t] == np.nanmax(win) and mn[t] == np.nanmin(win)


def test_window_nan_semantics():
    """A NaN affects exactly the windows containing it (IMPLEMENTATION_NOTES)."""
    o, h, l, c = DATASETS["rw0"]
    x = c.copy()
    x[100] = np.nan
    p = 20
    for f in (ORC.sma, ORC.wma, ORC.var, ORC.stddev):
        y = f(x, p)
        assert np.isnan(y[100: 100 + p]).all()
        assert not np.isnan(y[100 + p:]).any()
        assert not np.isnan(y[p - 1: 100]).any()


def test_recurrence_leading_nan_skip():
    """EMA/RSI/MACD/KAMA skip a leading NaN prefix (reference EMA.cu:65)."""
    o, h, l, c = DATASETS["rw0"]
    x = c.copy()
    x[:17] = np.nan
    for f, lb in ((lambda v: ORC.ema(v, 20), 19), (lambda v: ORC.rsi(v, 14), 14),
                  (lambda v: ORC.macd(v, 12, 26, 9)[0], 33), (lambda v: ORC.kama(v, 30), 30),
                  (lambda v: ORC.dema(v, 20), 38), (lambda v: ORC.tema(v, 20), 57),
                  (lambda v: ORC.t3(v, 5, 0.7), 24)):
        y = f(x)
        z = f(c[17:])
        assert np.isnan(y[: 17 + lb]).all()
        np.testing.assert_array_equal(y[17:], z)


def test_invalid_arguments():
    x = np.ones(10, np.float32)
    for bad in (lambda: ORC.ema(x, 0), lambda: ORC.sma(x, -2), lambda: ORC.t3(x, 1, 0.7),
                lambda: ORC.macd(x, 0, 26, 9)):
        with pytest.raises(ValueError):
            bad()
    assert np.isnan(ORC.ema(x, 20)).all()


def test_period_one_atr_and_natr_are_trange():
    """Reference contract (docs/BATCH_API.md:155-159): ATR(1) and NATR(1) are
    the unnormalized true range, NaN at bar 0, recovering after a NaN bar.
    On finite bars this is also TA-Lib 0.8.3's fma(0, prev, 1*TR)."""
    o, h, l, c = DATASETS["rw0"]
    h, l, c = h.copy(), l.copy(), c.copy()
    tr = ORC.trange(h, l, c)
    assert np.array_equal(ORC.atr(h, l, c, 1), tr, equal_nan=True)
    assert np.array_equal(ORC.natr(h, l, c, 1), tr, equal_nan=True)
    h[100] = np.nan
    tr = ORC.trange(h, l, c)
    atr = ORC.atr(h, l, c, 1)
    assert np.isnan(atr[100]) and np.isfinite(atr[101:]).all()
    assert np.array_equal(atr, tr, equal_nan=True)


def test_obv_and_ad_brute_force():
    """OBV / AD (reference OBV.cu / AD.cu semantics): numpy cumulative sums
    in float64, independent of the C loop."""
    import tacuda_amd as ta

    d = ta.synth_ohlcv(3000, seed=21)
    h, l, c, v = (d[k].astype(np.float64) for k in ("high", "low", "close", "volume"))
    sgn = np.sign(np.diff(c))
    obv = np.concatenate([[v[0]], v[0] + np.cumsum(sgn * v[1:])])
    assert np.allclose(ORC.obv(d["close"], d["volume"]), obv, rtol=1e-6)
    rng = h - l
    term = np.where(rng > 0, ((c - l) - (h - c)) / np.where(rng > 0, rng, 1) * v, 0.0)
    assert np.allclose(ORC.ad(d["high"], d["low"], d["close"], d["volume"]), np.cumsum(term),
                       rtol=1e-6, atol=1e-3)
    # non-finite bars contribute nothing to AD
    h2 = d["high"].copy()
    h2[100] = np.nan
    a1 = ORC.ad(h2, d["low"], d["close"], d["volume"])
    t2 = term.copy()
    t2[100] = 0.0
    assert np.allclose(a1, np.cumsum(t2), rtol=1e-6, atol=1e-3)


def test_host_aroon_two_periods_brute_force():
    """Host AROON with distinct periods: up and down each follow their own
    window (newest extremum wins ties), the oscillator is up - down from the
    longer period, and a period beyond the series gives three NaN lines."""
    o, h, l, c = DATASETS["rw1"]
    h, l = h[:400], l[:400]
    up_p, dn_p = 9, 25
    up, dn, osc = ORC.aroon_host(h, l, up_p, dn_p)
    assert np.isnan(up[:up_p]).all() and np.isnan(dn[:dn_p]).all()
    assert np.isnan(osc[:dn_p]).all()
    for t in range(up_p, len(h)):
        wh = h[t - up_p: t + 1][::-1]
        assert up[t] == np.float32(100.0 / up_p * (up_p - int(np.argmax(wh))))
    for t in range(dn_p, len(h)):
        wl = l[t - dn_p: t + 1][::-1]
        assert dn[t] == np.float32(100.0 / dn_p * (dn_p - int(np.argmin(wl))))
        assert osc[t] == np.float32(up[t] - dn[t])
    for line in ORC.aroon_host(h[:20], l[:20], 5, 21):
        assert np.isnan(line).all()


def test_engulfing_legacy_alias_bars():
    """The legacy engulfing names: +100 for a white body engulfing the black
    body before it, -100 for the mirror, +-80 when an open equals the other
    candle's close; the signal is written from bar 1 (the 61-pattern
    engulfing starts at bar 2) and equals it from bar 2 on."""
    # bar0 black 10->9, bar1 white 8.5->10.5 (engulfs): +100 at bar 1
    o = np.array([10.0, 8.5, 11.0, 10.0, 9.0, 9.0], np.float32)
    c = np.array([9.0, 10.5, 10.0, 11.5, 9.5, 8.0], np.float32)
    h = np.maximum(o, c) + 0.1
    l = np.minimum(o, c) - 0.1
    out = ORC.engulfing_legacy(o, h, l, c)
    assert out[0] == 0 and out[1] == 100
    # bar2 black 11->10 after white 8.5->10.5: open 11 > 10.5, close 10 > 8.5: no
    assert out[2] == 0
    # bar3 white 10->11.5 after black 11->10: open 10 == prev close -> +80
    assert out[3] == 80
    # bar5 black 9->8 after white 9->9.5: open 9 < 9.5 fails the strict rule,
    # but open >= ... no engulfing (close 8 < 9, open 9 < prev close 9.5)
    assert out[5] == 0
    full = ORC.candle("engulfing", o, h, l, c)
    np.testing.assert_array_equal(out[2:], full[2:].astype(np.float32))


@needs_ref
@pytest.mark.parametrize("ds", ["rw0", "rw1", "short", "zeros"])
@pytest.mark.parametrize("p", [2, 14, 30])
def test_min_max_pinned_by_donchian(ds, p):
    """MAX(high) and MIN(low) equal the vendored TA_DONCHIAN bands bit for
    bit (extrema are input values, so no rounding enters)."""
    o, h, l, c = DATASETS[ds]
    up, dn = REF.donchian(h, l, p)
    _bits_equal(ORC.max(h, p), up)
    _bits_equal(ORC.min(l, p), dn)


@needs_ref
@pytest.mark.parametrize("ds", ["rw0", "rw1", "short"])
@pytest.mark.parametrize("p", [5, 14, 30])
def test_willr_pinned_by_stochf(ds, p):
    """WILLR = fast %K - 100 (TA_STOCHF of the vendored library): the same
    NaN range and values within float rounding of the two forms."""
    o, h, l, c = DATASETS[ds]
    k, _ = REF.stochf(h, l, c, p, 1)
    w = ORC.willr(h, l, c, p)
    assert np.array_equal(np.isnan(w), np.isnan(k))
    ok = ~np.isnan(k)
    np.testing.assert_allclose(w[ok], k[ok].astype(np.float64) - 100.0, rtol=0, atol=3e-5)


@needs_ref
@pytest.mark.parametrize("ds", ["rw0", "rw1", "short", "zeros"])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 10, 11, 30])
def test_trima_matches_reference_talib(ds, p):
    o, h, l, c = DATASETS[ds]
    _bits_equal(ORC.trima(c, p), REF.trima(c, p))


MA_TYPES = [0, 1, 2, 3, 4, 5, 6, 8]  # TA_MAType without MAMA


@needs_ref
@pytest.mark.parametrize("ds", ["rw0", "rw1"])
@pytest.mark.parametrize("sgt", MA_TYPES)
def test_macdext_matches_reference_talib(ds, sgt):
    """Every fast/slow/signal moving-average combination (and reversed
    periods, and one-bar periods) against the vendored TA_MACDEXT: identical
    NaN ranges, values within 1e-5 (5e-5 with a KAMA signal line, whose
    efficiency ratio amplifies the float32 rounding of the macd line;
    corpus tolerance 1e-4)."""
    o, h, l, c = DATASETS[ds]
    tol = 5e-5 if sgt == 6 else 1e-5
    cases = [(12, ft, 26, st, 9, sgt) for ft in MA_TYPES for st in MA_TYPES]
    cases += [(26, 3, 12, 0, 9, sgt), (5, 8, 3, 6, 1, sgt), (2, 0, 7, 2, 4, sgt)]
    for args in cases:
        got, want = ORC.macdext(c, *args), REF.macdext(c, *args)
        for j, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(np.isnan(g), np.isnan(w)), (args, j)
            ok = ~np.isnan(w)
            rel = np.abs(g[ok].astype(np.float64) - w[ok]) / np.maximum(
                1.0, np.maximum(np.abs(g[ok]), np.abs(w[ok])))
            assert rel.size == 0 or rel.max() <= tol, (args, j, rel.max())


@needs_ref
def test_macdext_skips_a_leading_nan_prefix():
    """The formula backend convention: a NaN prefix is skipped and the rest
    is TA_MACDEXT of the remaining bars; one-bar fast or slow periods are
    rejected as TA_MACDEXT rejects them."""
    o, h, l, c = DATASETS["rw0"]
    x = c.copy()
    x[:37] = np.nan
    with pytest.raises(ValueError):
        ORC.macdext(c, 1, 0, 7, 0, 4, 0)
    with pytest.raises(ValueError):
        REF.macdext(c, 1, 0, 7, 0, 4, 0)
    got = ORC.macdext(x, 10, 2, 21, 4, 7, 0)
    want = REF.macdext(c[37:], 10, 2, 21, 4, 7, 0)
    for g, w in zip(got, want):
        assert np.isnan(g[:37]).all()
        ok = ~np.isnan(w)
        assert np.array_equal(np.isnan(g[37:]), np.isnan(w))
        np.testing.assert_allclose(g[37:][ok], w[ok], rtol=0, atol=1e-5)
