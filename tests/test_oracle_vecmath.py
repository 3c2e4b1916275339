"""Pins of the oracle's element-wise math, BOP, IBS, CHANGE, AVGDEV and MINMAX
(oracle/oracle.c): numpy's float32 operators for the exactly rounded
operations, numpy's float64 functions and special values for the
transcendental ones, the vendored TA-Lib 0.8.3 TA_IBS (oracle/_ref) bit for
bit, closed forms where float arithmetic is exact, and brute force over
short windows."""
import math

import numpy as np
import pytest

from oracle import ORC, REF, ref

needs_ref = pytest.mark.skipif(ref() is None, reason="reference TA-Lib host build not present")

NP1 = {"acos": np.arccos, "asin": np.arcsin, "atan": np.arctan, "ceil": np.ceil, "cos": np.cos,
       "cosh": np.cosh, "exp": np.exp, "floor": np.floor, "ln": np.log, "log10": np.log10,
       "sin": np.sin, "sinh": np.sinh, "sqrt": np.sqrt, "tan": np.tan, "tanh": np.tanh}
EXACT1 = ("ceil", "floor", "sqrt")


def _domain(name, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    if name in ("acos", "asin"):
        x = rng.uniform(-1, 1, n)
    elif name in ("ln", "log10", "sqrt"):
        x = np.exp(rng.uniform(-20, 20, n))
    elif name in ("cosh", "sinh", "exp"):
        x = rng.uniform(-80, 80, n)
    else:
        x = rng.uniform(-200, 200, n)
    return x.astype(np.float32)


def ulps(a, b):
    """Distance in float32 units in the last place (NaN/Inf positions must agree)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = np.isfinite(a) & np.isfinite(b)
    assert np.array_equal(a[~ok & ~np.isnan(a)], b[~ok & ~np.isnan(b)])

    def key(v):
        i = v.view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)

    return np.abs(key(a[ok]) - key(b[ok]))


@pytest.mark.parametrize("name", list(NP1))
def test_math1_against_numpy(name):
    x = _domain(name)
    got = ORC.math1(name, x)
    want = NP1[name](x.astype(np.float64)).astype(np.float32)
    if name in EXACT1:
        assert np.array_equal(got, want), name
        # and numpy's own float32 operation gives the same bits
        assert np.array_equal(got, NP1[name](x)), name
    else:
        assert ulps(got, want).max() <= 1, name


def test_math1_special_values():
    f = np.float32
    assert ORC.math1("exp", np.array([0.0], f))[0] == 1.0
    assert ORC.math1("ln", np.array([1.0], f))[0] == 0.0
    assert ORC.math1("log10", np.array([1000.0], f))[0] == 3.0
    assert ORC.math1("sqrt", np.array([4.0, 2.25], f)).tolist() == [2.0, 1.5]
    assert ORC.math1("ceil", np.array([-1.5, 1.25], f)).tolist() == [-1.0, 2.0]
    assert ORC.math1("floor", np.array([-1.5, 1.25], f)).tolist() == [-2.0, 1.0]
    assert ORC.math1("acos", np.array([1.0], f))[0] == 0.0
    assert ORC.math1("atan", np.array([np.inf], f))[0] == np.float32(math.pi / 2)
    assert np.isnan(ORC.math1("ln", np.array([-1.0], f))[0])
    assert np.isnan(ORC.math1("asin", np.array([2.0], f))[0])
    assert ORC.math1("ln", np.array([0.0], f))[0] == -np.inf
    # odd / even symmetry
    x = _domain("sin", 500)
    for name, sgn in (("sin", -1), ("tan", -1), ("sinh", -1), ("tanh", -1), ("atan", -1),
                      ("cos", 1), ("cosh", 1)):
        assert np.array_equal(ORC.math1(name, -x), sgn * ORC.math1(name, x)), name


@pytest.mark.parametrize("name", ["add", "sub", "mult", "div"])
def test_math2_bit_exact_against_numpy_float32(name):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000)).astype(np.float32)
    b = (rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000)).astype(np.float32)
    b[::97] = 0.0  # division by zero: signed infinities and NaN
    a[::389] = 0.0
    op = {"add": np.add, "sub": np.subtract, "mult": np.multiply, "div": np.divide}[name]
    with np.errstate(all="ignore"):
        want = op(a, b)
    got = ORC.math2(name, a, b)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32)), name


def _bits(a, b, name):
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{name}: NaN masks differ"
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), f"{name}: bits differ"


@needs_ref
@pytest.mark.parametrize("seed", [0, 3, 11])
def test_ibs_matches_vendored_talib(seed):
    import tacuda_amd as ta

    d = ta.synth_ohlcv(3000, seed=seed)
    h, l, c = d["high"].copy(), d["low"].copy(), d["close"].copy()
    h[10] = l[10]  # a zero range: 0.5
    h[20] = l[20] - 1.0  # a negative range: 0.5
    _bits(ORC.ibs(h, l, c), REF.ibs(h, l, c), f"ibs seed {seed}")


def test_ibs_and_bop_closed_forms():
    rng = np.random.default_rng(2)
    n = 600
    lo = rng.integers(1, 500, n).astype(np.float32)
    span = (2.0 ** rng.integers(0, 6, n)).astype(np.float32)  # power-of-two ranges
    hi = lo + span
    cl = lo + np.floor(rng.random(n) * span)
    op = lo + np.floor(rng.random(n) * span)
    assert np.array_equal(ORC.ibs(hi, lo, cl), ((cl - lo) / span).astype(np.float32))
    assert np.array_equal(ORC.bop(op, hi, lo, cl), ((cl - op) / span).astype(np.float32))
    # zero ranges
    z = lo.copy()
    assert np.array_equal(ORC.bop(op, z, lo, cl), np.zeros(n, np.float32))
    assert np.array_equal(ORC.ibs(z, lo, cl), np.full(n, 0.5, np.float32))


def test_change_against_numpy_float32():
    import tacuda_amd as ta

    x = ta.synth_ohlcv(2000, seed=4)["close"]
    for p in (1, 5, 60):
        got = ORC.change(x, p)
        assert np.isnan(got[:p]).all()
        assert np.array_equal(got[p:], x[p:] - x[:-p])
    assert np.isnan(ORC.change(x, 2000)).all()


def test_avgdev_closed_form_and_brute_force():
    # integer data and a power-of-two period: every float sum, the mean and
    # the deviations are exact, so the value is the exact mean deviation
    rng = np.random.default_rng(5)
    x = rng.integers(0, 1000, 700).astype(np.float32)
    for p in (1, 4, 16):
        w = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), p)
        want = np.abs(w - w.mean(1, keepdims=True)).mean(1).astype(np.float32)
        got = ORC.avgdev(x, p)
        assert np.isnan(got[:p - 1]).all()
        assert np.array_equal(got[p - 1:], want), p
    # float32 brute force (oldest first) on prices
    import tacuda_amd as ta

    y = ta.synth_ohlcv(300, seed=6)["close"]
    p = 7
    got = ORC.avgdev(y, p)
    for t in range(p - 1, 300):
        s = np.float32(0)
        for v in y[t - p + 1:t + 1]:
            s = np.float32(s + v)
        m = np.float32(s / np.float32(p))
        dv = np.float32(0)
        for v in y[t - p + 1:t + 1]:
            dv = np.float32(dv + np.float32(abs(np.float32(v - m))))
        assert got[t] == np.float32(dv / np.float32(p)), t


@pytest.mark.filterwarnings("ignore:All-NaN slice")
def test_minmax_against_numpy():
    import tacuda_amd as ta

    x = ta.synth_ohlcv(900, seed=7)["close"].copy()
    x[100] = np.nan
    x[400:410] = np.nan  # a run longer than p = 5: all-NaN windows
    for p in (1, 5, 30):
        mn, mx = ORC.minmax(x, p)
        w = np.lib.stride_tricks.sliding_window_view(x, p)
        assert np.isnan(mn[:p - 1]).all() and np.isnan(mx[:p - 1]).all()
        _bits(mn[p - 1:], np.nanmin(w, axis=1), f"min {p}")
        _bits(mx[p - 1:], np.nanmax(w, axis=1), f"max {p}")


def _rolling_brute(kind, x, p):
    """float64 restatement, newest first (independent loop)."""
    out = np.full(len(x), np.nan, np.float32)
    for t in range(p - 1, len(x)):
        w = x[t - p + 1:t + 1][::-1]
        if not np.all(np.isfinite(w)):
            continue
        s = 0.0
        q = 0.0
        for v in w.astype(np.float64):
            s += v
            q += v * v
        if kind == "mean":
            out[t] = np.float32(s / p)
        elif kind == "max":
            out[t] = w.max()
        elif kind == "min":
            out[t] = w.min()
        else:
            m = s / p
            out[t] = np.float32(np.sqrt(max(0.0, q / p - m * m)))
    return out


@pytest.mark.parametrize("kind", ["mean", "max", "min", "std"])
def test_rolling_against_brute_force(kind):
    import tacuda_amd as ta

    x = ta.synth_ohlcv(600, seed=12)["close"].copy()
    x[100] = np.nan
    x[300] = np.inf
    for p in (1, 5, 33):
        _bits(ORC.rolling(kind, x, p), _rolling_brute(kind, x, p), f"rolling {kind} {p}")
    with pytest.raises(ValueError):
        ORC.rolling(kind, x, 601)  # period beyond the series
    with pytest.raises(ValueError):
        ORC.rolling(kind, x, 0)


def test_rolling_closed_forms():
    c = np.full(50, 7.25, np.float32)
    assert np.all(ORC.rolling("std", c, 8)[7:] == 0.0)
    assert np.all(ORC.rolling("mean", c, 8)[7:] == 7.25)
    i = np.arange(1, 51, dtype=np.float32)  # integers: exact sums
    assert np.array_equal(ORC.rolling("mean", i, 4)[3:], (i[3:] - 1.5).astype(np.float32))
    assert np.array_equal(ORC.rolling("max", i, 4)[3:], i[3:])
    assert np.array_equal(ORC.rolling("min", i, 4)[3:], i[:-3])
    # population std of 4 consecutive integers: sqrt(1.25)
    assert np.allclose(ORC.rolling("std", i, 4)[3:], np.sqrt(1.25), rtol=1e-7)


def test_pivots_brute_force_and_shapes():
    import tacuda_amd as ta

    x = ta.synth_ohlcv(700, seed=14)["close"]
    for s in (1, 3, 8):
        for high in (True, False):
            want = np.full(len(x), np.nan, np.float32)
            for t in range(2 * s, len(x)):
                b = t - s
                left, right = x[b - s:b], x[b + 1:b + s + 1]
                if high and np.all(x[b] > left) and np.all(x[b] >= right):
                    want[t] = x[b]
                if not high and np.all(x[b] < left) and np.all(x[b] <= right):
                    want[t] = x[b]
            _bits(ORC.pivot(high, x, s), want, f"pivot high={high} s={s}")
    # a peak, a plateau (equal right neighbour counts, equal left one does not)
    y = np.array([1, 2, 5, 3, 1, 4, 4, 2, 0, 0], np.float32)
    hi = ORC.pivot(True, y, 1)
    assert hi[3] == 5.0 and hi[6] == 4.0 and np.isnan(hi[7])
    lo = ORC.pivot(False, y, 1)
    # a trough at bar 4; bar 8 (0 before an equal 0) counts: the right side may tie
    assert lo[5] == 1.0 and lo[9] == 0.0 and np.isnan(lo[8])
    with pytest.raises(ValueError):
        ORC.pivot(True, y, 5)  # 2 s >= size


def test_skew_against_numpy_and_closed_forms():
    import tacuda_amd as ta

    x = ta.synth_ohlcv(500, seed=15)["close"]
    for p in (3, 20):
        w = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), p)
        d = w - w.mean(1, keepdims=True)
        m2, m3 = (d ** 2).mean(1), (d ** 3).mean(1)
        want = np.where(m2 <= 0, 0.0, m3 / m2 ** 1.5)
        got = ORC.skew(x, p)
        assert np.isnan(got[:p - 1]).all()
        np.testing.assert_allclose(got[p - 1:], want, rtol=1e-4, atol=1e-6)
    sym = np.tile(np.array([1, 2, 3], np.float32), 20)  # symmetric windows
    assert np.allclose(ORC.skew(sym, 3)[2:], 0.0, atol=1e-7)
    assert np.all(ORC.skew(np.full(30, 4.0, np.float32), 5)[4:] == 0.0)
    with pytest.raises(ValueError):
        ORC.skew(sym, 61)


def test_trendscore_and_market_strength():
    import tacuda_amd as ta

    x = ta.synth_ohlcv(400, seed=16)["close"]
    for lb in (0, 5):
        want = np.full(400, np.nan, np.float32)
        for t in range(lb + 10, 400):
            want[t] = sum(1 if x[t] >= x[t - lb - k] else -1 for k in range(1, 11))
        _bits(ORC.trendscore(x, lb), want, f"trendscore {lb}")
    for p in (1, 9):
        want = np.full(400, np.nan, np.float32)
        d = np.diff(x.astype(np.float64))
        for t in range(p, 400):
            w = d[t - p:t]
            tot, ups, dns = w.sum(), w[w > 0].sum(), -w[w < 0].sum()
            want[t] = (100 * tot / ups if ups > 0 else 0.0) if tot >= 0 else \
                (100 * tot / dns if dns > 0 else 0.0)
        np.testing.assert_allclose(ORC.market_strength(x, p)[p:], want[p:], rtol=1e-6, atol=1e-5)
    up = np.arange(1, 60, dtype=np.float32)
    assert np.all(ORC.trendscore(up, 3)[13:] == 10.0)
    assert np.all(ORC.trendscore(up[::-1].copy(), 3)[13:] == -10.0)
    assert np.all(ORC.market_strength(up, 7)[7:] == 100.0)
    assert np.all(ORC.market_strength(up[::-1].copy(), 7)[7:] == -100.0)
    with pytest.raises(ValueError):
        ORC.trendscore(up, 49)  # size <= lookback + 10
