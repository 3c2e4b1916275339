"""Pins of the candlestick restatement (oracle/candles.c), CPU only: the
reference's hand-authored activation fixture, the signal set and warm-up of
every pattern, and independent numpy restatements of single- and two-candle
patterns (TA-Lib rules written directly over arrays)."""
import numpy as np
import pytest

from oracle import ORC, lib
from tests.candle_data import activation, quantized

NAMES = None


def names():
    global NAMES
    if NAMES is None:
        import ctypes

        L = lib()
        L.orc_cdl_name.restype = ctypes.c_char_p
        L.orc_cdl_lookback.restype = ctypes.c_int
        NAMES = [(L.orc_cdl_name(k).decode(), L.orc_cdl_lookback(k)) for k in range(L.orc_cdl_count())]
    return NAMES


def test_sixty_one_patterns_with_reference_names():
    assert len(names()) == 61
    assert len({n for n, _ in names()}) == 61


def test_activation_fixture_fires_where_documented():
    (o, h, l, c), fires = activation()
    for name, bar, sig in fires:
        out = ORC.candle(name, o, h, l, c)
        assert out[bar] == sig, (name, bar, out[bar])
    # the calm stretches between injections fire none of the rare patterns
    for name in ("three_black_crows", "identical_three_crows", "mat_hold", "stalled_pattern",
                 "conceal_baby_swallow", "three_stars_in_south"):
        out = ORC.candle(name, o, h, l, c)
        assert np.count_nonzero(out) == 1, name


@pytest.mark.parametrize("seed", [0, 1])
def test_signal_set_and_warmup(seed):
    o, h, l, c = quantized(4000, seed=seed)
    fired = 0
    for name, lb in names():
        out = ORC.candle(name, o, h, l, c)
        assert (out[:lb] == 0).all(), name
        vals = set(np.unique(out).tolist())
        allowed = {0, 100, -100}
        if name in ("engulfing", "harami", "harami_cross"):
            allowed |= {80, -80}
        if name in ("hikkake", "hikkake_mod"):
            allowed |= {200, -200}
        assert vals <= allowed, (name, vals)
        fired += bool(np.count_nonzero(out))
    assert fired >= 50  # most patterns occur on 4000 gapped, quantized bars


# ---- independent restatements (numpy, direct window means) -------------
def _avg(x, i, p):
    return x[i - p:i].mean() if p else x[i]


def _np_patterns(o, h, l, c):
    O, H, L, C = (a.astype(np.float64) for a in (o, h, l, c))
    rb = np.abs(C - O)
    hl = H - L
    us = H - np.maximum(O, C)
    ls = np.minimum(O, C) - L
    col = np.where(C >= O, 1, -1)
    n = len(C)
    res = {k: np.zeros(n, np.int32) for k in ("doji", "marubozu", "spinning_top",
                                              "engulfing", "long_line", "dragonfly_doji")}
    for i in range(n):
        if i >= 10:
            doji_thr = 0.1 * (hl[i - 10:i].sum() / 10)
            long_thr = 1.0 * (rb[i - 10:i].sum() / 10)
            vshort = 0.1 * (hl[i - 10:i].sum() / 10)
            sh = (us + ls)[i - 10:i].sum() / 10 / 2.0
            res["doji"][i] = 100 if rb[i] <= doji_thr else 0
            if rb[i] > long_thr and us[i] < vshort and ls[i] < vshort:
                res["marubozu"][i] = col[i] * 100
            if rb[i] < long_thr and us[i] > rb[i] and ls[i] > rb[i]:
                res["spinning_top"][i] = col[i] * 100
            if rb[i] > long_thr and us[i] < sh and ls[i] < sh:
                res["long_line"][i] = col[i] * 100
            if rb[i] <= doji_thr and us[i] < vshort and ls[i] > vshort:
                res["dragonfly_doji"][i] = 100
        if i >= 2:
            w = col[i] == 1 and col[i - 1] == -1 and (
                (C[i] >= O[i - 1] and O[i] < C[i - 1]) or (C[i] > O[i - 1] and O[i] <= C[i - 1]))
            b = col[i] == -1 and col[i - 1] == 1 and (
                (O[i] >= C[i - 1] and C[i] < O[i - 1]) or (O[i] > C[i - 1] and C[i] <= O[i - 1]))
            if w or b:
                strict = O[i] != C[i - 1] and C[i] != O[i - 1]
                res["engulfing"][i] = col[i] * (100 if strict else 80)
    return res


@pytest.mark.parametrize("seed", [0, 3])
def test_independent_numpy_restatement(seed):
    o, h, l, c = quantized(3000, seed=seed)
    want = _np_patterns(o, h, l, c)
    for name, w in want.items():
        got = ORC.candle(name, o, h, l, c)
        assert np.array_equal(got, w), (name, np.flatnonzero(got != w)[:5])
    assert np.any(np.abs(want["engulfing"]) == 80)  # the equal-end case occurs


def test_bullish_bearish_mirror():
    """Reflecting prices (x -> 2M - x, high <-> low) turns a bullish engulfing
    into a bearish one of the same strength, wherever no candle of the pair
    has close == open (TA-Lib colors those white either way)."""
    o, h, l, c = quantized(3000, seed=5)
    M = np.float32(200.0)
    ro, rh, rl, rc = M - o, M - l, M - h, M - c
    a = ORC.candle("engulfing", o, h, l, c)
    b = ORC.candle("engulfing", ro, rh, rl, rc)
    flat = c == o
    ok = ~(flat | np.roll(flat, 1))
    ok[0] = False
    assert np.array_equal(a[ok], -b[ok])
    assert np.count_nonzero(a[ok]) > 100
