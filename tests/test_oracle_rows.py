"""Pins of the oracle's TA-Lib 0.8.3 row functions (oracle/oracle.c
orc_rows): bit-identical to the reference's own host build of TA-Lib 0.8.3
(oracle/_ref ref_rows, which wraps TA_CUMSUM ... TA_RVOL in the backend's
leading-NaN rule) on seeded bars, periods at and around the edges of
TA-Lib's ranges, zero-volume and unchanged-price runs, leading and interior
NaNs; plus closed forms that fix a few of them independently of TA-Lib."""
import numpy as np
import pytest

from oracle import ORC, REF, ref

needs_ref = pytest.mark.skipif(ref() is None, reason="reference TA-Lib host build not present")

# inputs by letter: x close, o open, h high, l low, c close, v volume
SPEC = {"cumsum": "x", "pvt": "cv", "rma": "x", "zlema": "x", "efi": "cv", "vwma": "xv",
        "er": "x", "vhf": "x", "dpo": "x", "chop": "hlc", "adr": "hl", "ao": "hl",
        "qstick": "oc", "donchian": "hl", "marketfi": "hlv", "wad": "hlc", "mcgd": "x",
        "rvol": "v", "cmf": "hlcv", "emv": "hlv", "fosc": "x", "eri": "hlc", "percentrank": "x",
        "si": "ohlc", "asi_wilder": "ohlc", "vortex": "hlc", "cmou": "x"}
# integer periods: 0, 1 or 2; "d": one real parameter; "1d": a period and a real
NPAR = {"cumsum": 0, "pvt": 0, "marketfi": 0, "wad": 0, "ao": 2, "emv": "1d", "si": "d",
        "asi_wilder": "d"}
PERIODS = {1: [(1, 0), (2, 0), (14, 0), (57, 0)], 2: [(2, 3), (5, 34), (34, 5), (9, 9)],
           0: [(0, 0)], "d": [(0, 0, 3.0), (0, 0, 0.5), (0, 0, 1e4)],
           "1d": [(1, 0, 1.0), (2, 0, 1e4), (14, 0, 1e8), (57, 0, 10.0)]}
MIN_P = {"er": 2, "vhf": 2, "dpo": 2, "chop": 2, "donchian": 2, "mcgd": 2, "cmf": 2, "fosc": 2,
         "percentrank": 2, "cmou": 2}


def bars(n, seed, holes=False):
    import tacuda_amd as ta

    d = {k: np.asarray(v, np.float32).copy() for k, v in ta.synth_ohlcv(n, seed=seed).items()}
    if n > 300:
        d["volume"][50:80] = 0.0  # a zero-volume run longer than the periods tested
        d["close"][120:140] = d["close"][119]  # unchanged closes
        d["volume"][200] = 0.0
    if holes:
        for k in d:
            d[k][:7] = np.nan  # leading prefix common to the inputs
        d["high"][7] = np.nan  # one more bar where one input is NaN
        d["close"][300] = np.nan  # an interior NaN
    return d


def inputs(name, d):
    key = {"x": "close", "o": "open", "h": "high", "l": "low", "c": "close", "v": "volume"}
    return [d[key[k]] for k in SPEC[name]]


def params(name):
    k = NPAR.get(name, 1)
    return [p for p in PERIODS[k] if k not in (1, "1d") or p[0] >= MIN_P.get(name, 1)]


def bits_equal(a, b, what):
    a = a if isinstance(a, tuple) else (a,)
    b = b if isinstance(b, tuple) else (b,)
    for k, (g, w) in enumerate(zip(a, b)):
        assert np.array_equal(np.isnan(g), np.isnan(w)), f"{what} out {k}: NaN masks differ"
        ok = ~np.isnan(g)
        assert np.array_equal(g[ok].view(np.uint32), w[ok].view(np.uint32)), \
            f"{what} out {k}: bits differ at {np.flatnonzero(g[ok] != w[ok])[:5]}"


@needs_ref
@pytest.mark.parametrize("name", list(SPEC))
@pytest.mark.parametrize("seed,holes", [(0, False), (3, False), (11, True)])
def test_rows_match_vendored_talib(name, seed, holes):
    d = bars(2500, seed, holes)
    for p in params(name):
        bits_equal(ORC.rows(name, inputs(name, d), *p), REF.rows(name, inputs(name, d), *p),
                   f"{name}{p} seed {seed}")


@needs_ref
@pytest.mark.parametrize("name", [n for n in SPEC if NPAR.get(n, 1) == 1])
def test_rows_short_series_and_invalid_periods(name):
    hi = 10000 if name == "percentrank" else 100000
    d = bars(40, 1)
    lo = MIN_P.get(name, 1)
    for p in (lo, 39, 40, 41, 1000):  # around the series length
        bits_equal(ORC.rows(name, inputs(name, d), p), REF.rows(name, inputs(name, d), p),
                   f"{name} p{p}")
    for bad in (lo - 1, hi + 1):
        with pytest.raises(ValueError):
            ORC.rows(name, inputs(name, d), bad)
        with pytest.raises(ValueError):
            REF.rows(name, inputs(name, d), bad)


@needs_ref
def test_rows_all_nan_row_keeps_the_last_bar():
    d = bars(30, 2)
    for k in d:
        d[k][:] = np.nan
    for name in SPEC:
        p = params(name)[0]
        bits_equal(ORC.rows(name, inputs(name, d), *p), REF.rows(name, inputs(name, d), *p), name)


@needs_ref
def test_rows_real_parameter_ranges():
    d = bars(60, 4)
    for name, good, bad in (("emv", (5, 0, 1.0), (5, 0, 0.999)), ("emv", (5, 0, 3e37), (5, 0, 4e37)),
                            ("si", (0, 0, 1e-8), (0, 0, 5e-9)),
                            ("asi_wilder", (0, 0, 2.0), (0, 0, float("nan")))):
        bits_equal(ORC.rows(name, inputs(name, d), *good), REF.rows(name, inputs(name, d), *good),
                   f"{name}{good}")
        with pytest.raises(ValueError):
            ORC.rows(name, inputs(name, d), *bad)
        with pytest.raises(ValueError):
            REF.rows(name, inputs(name, d), *bad)


def test_rows_closed_forms():
    n = 64
    x = np.arange(1, n + 1, dtype=np.float32)  # 1, 2, ..., n
    # CUMSUM of 1..t is t(t+1)/2
    assert np.array_equal(ORC.rows("cumsum", [x]), (x * (x + 1) / 2).astype(np.float32))
    # DPO of a line: x[t - (p/2+1)] - mean(x[t-p+1..t]) = -(p/2+1) + (p-1)/2
    p = 6
    want = np.float32(-(p // 2 + 1) + (p - 1) / 2)
    assert np.all(ORC.rows("dpo", [x], p)[p - 1:] == want)
    # ER of a monotone series is 1; VHF of a line: the extremes span the
    # window's last p bars (p - 1) and the changes all p + 1 bars (p)
    assert np.all(ORC.rows("er", [x], 5)[5:] == 1.0)
    assert np.all(ORC.rows("vhf", [x], 5)[5:] == np.float32(4 / 5))
    # RMA of a constant is the constant; ADR of a constant range is the range
    c = np.full(n, 3.5, np.float32)
    assert np.all(ORC.rows("rma", [c], 7)[6:] == 3.5)
    assert np.all(ORC.rows("adr", [c + 2, c], 5)[4:] == 2.0)
    # DONCHIAN of a line: upper x[t], lower x[t-p+1], middle their mean
    u, m, lo = ORC.rows("donchian", [x, x], 4)
    assert np.array_equal(u[3:], x[3:]) and np.array_equal(lo[3:], x[:-3])
    assert np.array_equal(m[3:], (x[3:] + x[:-3]) / 2)
    # MARKETFI: (h - l) / v, 0 on zero volume
    v = np.array([2, 0, 4], np.float32)
    assert ORC.rows("marketfi", [np.array([5, 5, 9], np.float32),
                                 np.array([1, 1, 1], np.float32), v]).tolist() == [2.0, 0.0, 2.0]
    # PERCENTRANK of a rising series is 100, of a falling one 0
    assert np.all(ORC.rows("percentrank", [x], 5)[5:] == 100.0)
    assert np.all(ORC.rows("percentrank", [x[::-1].copy()], 5)[5:] == 0.0)
    # ERI of constant bars: bull = high - close, bear = low - close
    hh, ll, cc = np.full(n, 5.0, np.float32), np.full(n, 1.0, np.float32), c
    bull, bear = ORC.rows("eri", [hh, ll, cc], 4)
    assert np.all(bull[3:] == 1.5) and np.all(bear[3:] == -2.5)
    # CMF with the close at the high on every bar: money-flow multiplier 1
    vv = np.full(n, 7.0, np.float32)
    assert np.all(ORC.rows("cmf", [hh, ll, hh, vv], 5)[4:] == 1.0)
