"""Pins of the oracle's price transforms, momentum family, window midpoints
and linear-regression family (oracle/oracle.c): bit-identical to the
reference's host build of TA-Lib 0.8.3 (oracle/_ref) for the functions it
vendors (ROC, ROCP, MEDPRICE, TYPPRICE, MIDPRICE, LINEARREG), and for the
others checks that fix the result independently of the restatement: exact
closed forms on data where float arithmetic is exact, numpy brute force,
and identities between the family members."""
import math

import numpy as np
import pytest

from oracle import ORC, REF, ref

DATA_SEEDS = (0, 3, 11)


def _data(seed, n=2500):
    import tacuda_amd as ta

    return ta.synth_ohlcv(n, seed=seed)


def _bits(a, b, name):
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{name}: NaN masks differ"
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), f"{name}: bits differ"


needs_ref = pytest.mark.skipif(ref() is None, reason="reference TA-Lib host build not present")


@needs_ref
@pytest.mark.parametrize("seed", DATA_SEEDS)
@pytest.mark.parametrize("p", [1, 2, 10, 37])
def test_roc_rocp_match_vendored_talib(seed, p):
    d = _data(seed)
    x = d["close"].copy()
    x[100] = 0.0  # a zero earlier price: TA-Lib's 0 output
    _bits(ORC.roc(x, p), REF.roc(x, p), f"roc{p}")
    _bits(ORC.rocp(x, p), REF.rocp(x, p), f"rocp{p}")


@needs_ref
@pytest.mark.parametrize("seed", DATA_SEEDS)
def test_price_transforms_match_vendored_talib(seed):
    d = _data(seed)
    _bits(ORC.medprice(d["high"], d["low"]), REF.medprice(d["high"], d["low"]), "medprice")
    _bits(ORC.typprice(d["high"], d["low"], d["close"]),
          REF.typprice(d["high"], d["low"], d["close"]), "typprice")


@needs_ref
@pytest.mark.parametrize("seed", DATA_SEEDS)
@pytest.mark.parametrize("p", [2, 5, 14, 60])
def test_midprice_and_linearreg_match_vendored_talib(seed, p):
    d = _data(seed)
    _bits(ORC.midprice(d["high"], d["low"], p), REF.midprice(d["high"], d["low"], p), "midprice")
    _bits(ORC.linearreg(d["close"], p), REF.linearreg(d["close"], p), "linearreg")


def test_momentum_family_closed_forms():
    # x_t = 2^(t mod 8) * 3: every ratio is a power of two, every difference
    # an exact integer, so the float results are known exactly
    n, p = 200, 5
    t = np.arange(n)
    x = (3.0 * 2.0 ** (t % 8)).astype(np.float32)
    prev = np.roll(x, p).astype(np.float64)
    want_mom = (x - prev).astype(np.float32)
    want_rocr = (x / prev).astype(np.float32)
    for name, got, want in (("momentum", ORC.momentum(x, p), want_mom),
                            ("rocr", ORC.rocr(x, p), want_rocr),
                            ("rocr100", ORC.rocr100(x, p), (want_rocr * 100).astype(np.float32))):
        assert np.isnan(got[:p]).all(), name
        assert np.array_equal(got[p:], want[p:]), name
    # a zero earlier price gives 0 for the ratios, the difference for MOMENTUM
    z = x.copy()
    z[10] = 0.0
    assert ORC.rocr(z, p)[10 + p] == 0.0 and ORC.rocr100(z, p)[10 + p] == 0.0
    assert ORC.momentum(z, p)[10 + p] == x[10 + p]
    # ROCP = ROCR - 1 and ROC = 100 (ROCR - 1) on these exact ratios
    assert np.array_equal(ORC.rocp(x, p)[p:], (want_rocr - 1)[p:])
    assert np.array_equal(ORC.roc(x, p)[p:], ((want_rocr - 1) * 100).astype(np.float32)[p:])


def test_momentum_family_against_numpy():
    d = _data(5, 1000)
    x = d["close"]
    for p in (1, 7, 30):
        prev = x[:-p].astype(np.float64)
        cur = x[p:].astype(np.float64)
        np.testing.assert_array_equal(ORC.momentum(x, p)[p:], (cur - prev).astype(np.float32))
        np.testing.assert_array_equal(ORC.rocr(x, p)[p:], (cur / prev).astype(np.float32))
        np.testing.assert_array_equal(ORC.rocr100(x, p)[p:],
                                      ((cur / prev) * 100.0).astype(np.float32))
    assert np.isnan(ORC.momentum(x, 1000)).all()  # period >= size: all NaN


def test_avgprice_wclprice_closed_forms():
    rng = np.random.default_rng(4)
    o, h, l, c = (rng.integers(1, 1000, 500).astype(np.float32) for _ in range(4))
    # integer prices: the sums are exact and the quarters exact binary fractions
    assert np.array_equal(ORC.avgprice(o, h, l, c),
                          ((o.astype(np.float64) + h + l + c) / 4).astype(np.float32))
    assert np.array_equal(ORC.wclprice(h, l, c),
                          ((h.astype(np.float64) + l + 2 * c.astype(np.float64)) / 4)
                          .astype(np.float32))
    # on equal inputs every transform is the input
    x = rng.random(300).astype(np.float32) * 100
    for got in (ORC.avgprice(x, x, x, x), ORC.wclprice(x, x, x), ORC.medprice(x, x)):
        assert np.array_equal(got, x)


def _sliding(fn, x, p):
    out = np.full(len(x), np.nan, np.float32)
    for t in range(p - 1, len(x)):
        out[t] = fn(x[t - p + 1:t + 1])
    return out


@pytest.mark.filterwarnings("ignore:All-NaN slice")
def test_midpoint_against_numpy_and_midprice():
    d = _data(7, 800)
    x = d["close"].copy()
    x[300] = np.nan  # an interior NaN is skipped
    for p in (1, 4, 21):
        want = _sliding(lambda w: np.float32((np.float64(np.nanmax(w)) + np.nanmin(w)) / 2), x, p)
        _bits(ORC.midpoint(x, p), want, f"midpoint{p}")
        y = d["close"]
        _bits(ORC.midpoint(y, p), ORC.midprice(y, y, p), f"midpoint=midprice {p}")
    # a window starting inside the leading NaN run is NaN
    z = d["close"].copy()
    z[:5] = np.nan
    got = ORC.midpoint(z, 3)
    assert np.isnan(got[:7]).all() and not np.isnan(got[7])


def test_linear_regression_family_on_exact_lines():
    # y_t = a + b t with small integers: the regression recovers the line
    a, b, n = 7.0, 3.0, 300
    t = np.arange(n, dtype=np.float64)
    x = (a + b * t).astype(np.float32)
    for p in (2, 5, 14):
        s = np.arange(p - 1, n)
        assert np.array_equal(ORC.linearreg_slope(x, p)[s], np.full(len(s), b, np.float32))
        assert np.array_equal(ORC.linearreg(x, p)[s], (a + b * s).astype(np.float32))
        assert np.array_equal(ORC.linearreg_intercept(x, p)[s],
                              (a + b * (s - p + 1)).astype(np.float32))
        assert np.array_equal(ORC.tsf(x, p)[s], (a + b * (s + 1)).astype(np.float32))
        want_angle = np.float32(math.atan(b) * 180.0 / math.pi)
        assert np.allclose(ORC.linearreg_angle(x, p)[s], want_angle, rtol=0, atol=1e-5)
        assert np.isnan(ORC.linearreg(x, p)[:p - 1]).all()
    # one-bar windows: slope 0, intercept and forecasts the bar itself
    assert np.array_equal(ORC.linearreg_slope(x, 1), np.zeros(n, np.float32))
    assert np.array_equal(ORC.linearreg(x, 1), x) and np.array_equal(ORC.tsf(x, 1), x)


def test_linear_regression_identities_on_random_data():
    d = _data(9, 1500)
    x = d["close"]
    for p in (3, 14, 50):
        m = ORC.linearreg_slope(x, p).astype(np.float64)
        b = ORC.linearreg_intercept(x, p).astype(np.float64)
        ok = ~np.isnan(m)
        lr = ORC.linearreg(x, p).astype(np.float64)
        tsf = ORC.tsf(x, p).astype(np.float64)
        np.testing.assert_allclose(lr[ok], (b + (p - 1) * m)[ok], rtol=1e-6)
        np.testing.assert_allclose(tsf[ok], (b + p * m)[ok], rtol=1e-6)
        # the slope of an ordinary least-squares fit (numpy) on each window
        w = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), p)
        tt = np.arange(p)
        slope = ((tt - tt.mean()) * (w - w.mean(1, keepdims=True))).sum(1) / \
            ((tt - tt.mean()) ** 2).sum()
        np.testing.assert_allclose(m[p - 1:], slope, rtol=1e-6, atol=1e-9)
