"""Constructive pins of the candlestick oracle (oracle/candles.c), CPU only.

For each of the 45 patterns that neither the reference's activation fixture
nor the numpy restatements in test_oracle_candles.py pin by value, a
hand-built bar sequence that must fire with a stated signal at its last bar,
and near-miss variants in which exactly one condition of the TA-Lib rule is
broken by one price tick (or met with equality where the rule is strict),
which must not fire.  The sequences were derived from TA-Lib's rule
statements (as the reference restates them in
include/utils/CandleUtils.h:1510-2339), not from the oracle's text.

Every sequence starts with 15 identical baseline candles, open 100, high
101.5, low 99.5, close 101 (white body 1, shadows 0.5 and 0.5, range 2).
Averages over baseline windows are therefore: BodyLong = BodyShort = 1,
BodyDoji = 0.1 x 2 = 0.2, ShadowVeryShort = 0.2, ShadowShort = 1 / 2 = 0.5,
Near = 0.2 x 2 = 0.4, Far = 1.2, Equal = 0.1; ShadowLong and ShadowVeryLong
are 1 and 2 times the candle's own body.  Where a pattern's earlier candles
enter a later candle's window, the comment gives the shifted average.  All
prices are multiples of 1/16, exact in float32.
"""
import numpy as np
import pytest

from oracle import ORC

BASE = (100.0, 101.5, 99.5, 101.0)


def run(name, candles, nbase=15):
    rows = [BASE] * nbase + list(candles)
    o, h, l, c = (np.array(col, np.float32) for col in zip(*rows))
    assert (h >= np.maximum(o, c)).all() and (l <= np.minimum(o, c)).all(), "bad candle"
    return ORC.candle(name, o, h, l, c)


# name -> (candles, signal at the last bar, [near-miss candle lists])
LONG_BLACK = (103.0, 103.25, 99.75, 100.0)   # body 3, low 99.75, close 100
LONG_WHITE = (100.0, 103.25, 99.75, 103.0)   # body 3, high 103.25, close 103

CASES = {
    # RB > BodyLong(1), white, lower shadow < ShadowVeryShort(0.2)
    "belt_hold": ([(100, 103.25, 100, 103)], 100, [
        [(100, 103.25, 99.75, 103)],        # lower shadow 0.25
        [(100, 101.25, 100, 101)],          # body 1 = BodyLong, not above
    ]),
    # RB > BodyLong, white, upper shadow < 0.2
    "closing_marubozu": ([(100, 103, 99, 103)], 100, [
        [(100, 103.25, 99, 103)],           # upper shadow 0.25
    ]),
    # RB < BodyShort(1), both shadows > 2 x RB
    "high_wave": ([(100, 102, 98, 100.5)], 100, [
        [(100, 101.5, 98, 100.5)],          # upper shadow 1.0 = 2 x RB, not above
        [(100, 102, 98, 101)],              # body 1 = BodyShort, not below
    ]),
    # RB <= BodyDoji(0.2) and one shadow > RB
    "long_legged_doji": ([(100, 101, 99, 100.125)], 100, [
        [(100, 101, 99, 100.25)],           # body 0.25 > 0.2
        [(100, 100.125, 100, 100.125)],     # no shadow longer than the body
    ]),
    # RB < BodyShort, both shadows < ShadowShort(0.5)
    "short_line": ([(100, 100.75, 99.75, 100.5)], 100, [
        [(100, 101, 99.75, 100.5)],         # upper shadow 0.5, not below
    ]),
    # doji, both shadows > RB, body within Near(0.4) of the range midpoint
    "rickshaw_man": ([(100, 101, 99, 100.125)], 100, [
        [(100.5, 101, 99, 100.625)],        # body low 100.5 > midpoint 100 + 0.4
    ]),
    # doji, upper shadow < 0.2, lower shadow > 2 x RB
    "takuri": ([(100, 100.125, 98, 100.125)], 100, [
        [(100, 100.375, 98, 100.125)],      # upper shadow 0.25
    ]),
    # black long, then black body <= BodyShort (1.2 with the long body in the
    # window) opening below and closing above the first
    "homing_pigeon": ([LONG_BLACK, (102, 102.25, 100.75, 101)], 100, [
        [LONG_BLACK, (103, 103.25, 101.75, 102)],  # open 103 = open[i-1], not below
    ]),
    # two black closes within Equal(0.1)
    "matching_low": ([LONG_BLACK, (102, 102.25, 99.75, 100.0625)], 100, [
        [LONG_BLACK, (102, 102.25, 99.75, 100.125)],  # closes 0.125 apart
    ]),
    # long black, white opening below its low and closing within Equal of it
    "on_neck": ([LONG_BLACK, (99, 99.875, 98.75, 99.75)], -100, [
        [LONG_BLACK, (99, 99.875, 98.75, 99.875)],  # close 0.125 above the low
    ]),
    # long black, white opening below its low, closing above close+Equal and
    # at most at the black body's midpoint (101.5)
    "thrusting": ([LONG_BLACK, (99.5, 101.25, 99.25, 101.0)], -100, [
        [LONG_BLACK, (99.5, 101.75, 99.25, 101.625)],  # close above the midpoint
        [LONG_BLACK, (99.75, 101.25, 99.25, 101.0)],   # open = low[i-1], not below
    ]),
    # opposite colors, same open within Equal, long white (BodyLong 1.0 with
    # the black body 1 in the window) with lower shadow < ShadowVeryShort
    # (0.195)
    "separating_lines": ([(103, 103.25, 101.75, 102), (103, 106, 103, 105.75)], 100, [
        [(103, 103.25, 101.75, 102), (103.125, 106, 103.125, 105.75)],  # opens 0.125 apart
    ]),
    # long black then long white (BodyLong 1.2) closing at the same level
    "counterattack": ([LONG_BLACK, (97, 100.25, 96.75, 100)], 100, [
        [LONG_BLACK, (97, 100.25, 96.75, 100.125)],  # closes 0.125 apart
    ]),
    # two white candles gapping up over the one before, similar bodies
    # (within Near 0.38), opens within Equal (0.095)
    "gap_side_side_white": ([(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103),
                             (102, 103.25, 101.75, 103.125)], 100, [
        [(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103),
         (102.125, 103.5, 101.75, 103.25)],  # opens 0.125 apart
        [(100, 101.25, 99.75, 101), (101, 103.25, 100.75, 102),
         (101, 103.25, 100.75, 102.125)],    # second body does not gap
    ]),
    # long white, small black gapping up, black opening above and closing
    # inside the gap's black body but above the first close
    "upside_gap_2_crows": ([LONG_WHITE, (104.5, 104.75, 103.75, 104),
                            (105, 105.25, 103.25, 103.5)], -100, [
        [LONG_WHITE, (104.5, 104.75, 103.75, 104), (105, 105.25, 102.75, 103)],  # close = 103
    ]),
    # long black, small body inside it, white closing above the first open
    "three_inside": ([LONG_BLACK, (101, 101.75, 100.75, 101.5),
                      (101.5, 103.75, 101.25, 103.5)], 100, [
        [LONG_BLACK, (101, 101.75, 100.75, 101.5), (101.5, 103.25, 101.25, 103)],  # close = open[i-2]
        [LONG_BLACK, (101, 103.25, 100.75, 103), (101.5, 103.75, 101.25, 103.5)],  # 2nd body reaches 103
    ]),
    # three dojis (<= BodyDoji of the first, 0.2); the second gaps down, the
    # third's body low is above the second's
    "tristar": ([(100, 100.5, 99.5, 100.125), (99, 99.5, 98.5, 99.125),
                 (99.5, 100, 99, 99.625)], 100, [
        [(100, 100.5, 99.5, 100.125), (99, 99.5, 98.5, 99.125), (99, 99.5, 98.5, 99.125)],
        [(100, 100.5, 99.5, 100.125), (99, 99.5, 98.5, 99.125), (99.5, 100, 99, 99.75)],  # body 0.25
    ]),
    # long white, doji whose range gaps above it, black whose range gaps
    # below the doji, closing below close[i-2] - 0.3 x 3 = 102.1; BodyShort
    # at the third candle is (8 + 3 + 0.0625) / 10
    "abandoned_baby": ([LONG_WHITE, (104, 104.5, 103.5, 104.0625),
                        (103.375, 103.375, 101.5, 102.0)], -100, [
        [LONG_WHITE, (104, 104.5, 103.5, 104.0625), (103.375, 103.375, 101.75, 102.125)],
        [LONG_WHITE, (104, 104.5, 103.25, 104.0625), (103.375, 103.375, 101.5, 102.0)],  # no range gap
    ]),
    # three rising whites, the last with a shrinking body and a long upper
    # shadow
    "advance_block": ([(100, 103.125, 99.75, 103), (102, 105.125, 101.75, 105),
                       (104, 107, 103.75, 105.25)], -100, [
        [(100, 103.125, 99.75, 103), (102, 105.125, 101.75, 105), (104, 107, 103.75, 105)],
        [(100, 103.75, 99.75, 103), (102, 105.125, 101.75, 105), (104, 107, 103.75, 105.25)],
    ]),
    # long black, black gapping down, two lower candles, white closing into
    # the gap (above open[i-3] 102, below close[i-4] 103)
    "breakaway": ([(106, 106.25, 102.75, 103), (102, 102.25, 100.75, 101),
                   (101, 101.75, 99.75, 100), (100, 100.5, 98.75, 99),
                   (99, 102.75, 98.75, 102.5)], 100, [
        [(106, 106.25, 102.75, 103), (102, 102.25, 100.75, 101), (101, 101.75, 99.75, 100),
         (100, 100.5, 98.75, 99), (99, 103.25, 98.75, 103)],   # close = close[i-4]
    ]),
    # long white, black opening above its high and closing below the body's
    # midpoint (101.5) but above its open
    "dark_cloud_cover": ([LONG_WHITE, (103.5, 103.75, 100.5, 101.25)], -100, [
        [LONG_WHITE, (103.5, 103.75, 100.5, 101.5)],   # close = midpoint, not below
        [LONG_WHITE, (103.25, 103.5, 100.5, 101.25)],  # open = high[i-1], not above
    ]),
    # long white then a doji (<= 0.215) whose body gaps above it
    "doji_star": ([LONG_WHITE, (103.5, 104, 103.25, 103.5625)], -100, [
        [LONG_WHITE, (103, 103.5, 102.75, 103.0625)],  # body low 103 = close[i-1]
    ]),
    # doji, lower shadow < 0.2, upper shadow > 0.2
    "gravestone_doji": ([(100, 102, 100, 100.0625)], 100, [
        [(100, 102, 99.75, 100.0625)],      # lower shadow 0.25
    ]),
    # long white, doji gapping up, black (body > 1.10625) closing below 102.1
    "evening_doji_star": ([LONG_WHITE, (104, 104.5, 103.5, 104.0625),
                           (103.5, 103.75, 101.75, 102)], -100, [
        [LONG_WHITE, (104, 104.5, 103.5, 104.0625), (103.5, 103.75, 101.75, 102.125)],
        [LONG_WHITE, (104, 104.5, 103.5, 104.25), (103.5, 103.75, 101.75, 102)],  # star body 0.25
    ]),
    # the same with a small (<= 1.2) black star
    "evening_star": ([LONG_WHITE, (104, 104.75, 103.75, 104.5),
                      (103.5, 103.75, 101.75, 102)], -100, [
        [LONG_WHITE, (104, 104.75, 103.75, 104.5), (103.5, 103.75, 101.75, 102.125)],
    ]),
    # mirror: long black, doji gapping down, white closing above 100.9
    "morning_doji_star": ([LONG_BLACK, (99, 99.5, 98.5, 98.9375),
                           (99.5, 101.25, 99.25, 101)], 100, [
        [LONG_BLACK, (99, 99.5, 98.5, 98.9375), (99.5, 101.25, 99.25, 100.875)],
    ]),
    "morning_star": ([LONG_BLACK, (99, 99.25, 98.25, 98.5), (99.5, 101.25, 99.25, 101)], 100, [
        [LONG_BLACK, (99, 99.25, 98.25, 98.5), (99.5, 101.25, 99.25, 100.875)],
        [LONG_BLACK, (100, 100.25, 98.25, 98.5), (99.5, 101.25, 99.25, 101)],  # star top 100: no gap
    ]),
    # long black, small body strictly inside it (<= BodyShort 1.2)
    "harami": ([LONG_BLACK, (101, 101.75, 100.75, 101.5)], 100, [
        [LONG_BLACK, (102.5, 103.25, 102.25, 103.125)],  # body top 103.125 > 103
    ]),
    # long black, doji (<= 0.215) strictly inside
    "harami_cross": ([LONG_BLACK, (101, 101.5, 100.5, 101.0625)], 100, [
        [LONG_BLACK, (101, 101.5, 100.5, 101.25)],       # body 0.25
    ]),
    # small body, lower shadow > RB, upper shadow < 0.2, body low within
    # Near (0.4) above the previous low 99.5
    "hammer": ([(99.75, 100.0, 98.5, 100.0)], 100, [
        [(100.0, 100.25, 98.5, 100.25)],    # body low 100 > 99.9
    ]),
    # the same shape with its body low at or above previous high - Near (101.1)
    "hanging_man": ([(101.25, 101.5, 100, 101.5)], -100, [
        [(101.0, 101.25, 99.75, 101.25)],   # body low 101 < 101.1
    ]),
    # small body, upper shadow > RB, lower shadow < 0.2, body gapping below
    # the previous body (100..101)
    "inverted_hammer": ([(99.5, 101, 99.5, 99.75)], 100, [
        [(99.75, 101.25, 99.75, 100.0)],    # body top 100 = previous body low
    ]),
    "shooting_star": ([(101.5, 103, 101.5, 101.75)], -100, [
        [(101.0, 102.5, 101.0, 101.25)],    # body low 101 = previous body top
    ]),
    # long black, white opening below its low, closing within Equal above
    # the black close
    "in_neck": ([LONG_BLACK, (99.5, 100.25, 99.25, 100.0625)], -100, [
        [LONG_BLACK, (99.5, 100.25, 99.25, 100.125)],   # 0.125 above
        [LONG_BLACK, (99.5, 100.25, 99.25, 99.9375)],   # below the black close
    ]),
    # three falling blacks, a black with an upper shadow, a white opening
    # above it and closing above its high (103)
    "ladder_bottom": ([(105, 105.25, 103.75, 104), (104.5, 104.75, 102.75, 103),
                       (103.5, 103.75, 101.75, 102), (102, 103, 100.75, 101),
                       (102.5, 103.5, 102.25, 103.25)], 100, [
        [(105, 105.25, 103.75, 104), (104.5, 104.75, 102.75, 103), (103.5, 103.75, 101.75, 102),
         (102, 103, 100.75, 101), (102.5, 103.25, 102.25, 103.0)],   # close = high[i-1]
    ]),
    # long black, long white (> 1.2) opening below its low and closing above
    # its body midpoint (101.5) but below its open
    "piercing": ([LONG_BLACK, (99.5, 101.75, 99.25, 101.625)], 100, [
        [LONG_BLACK, (99.5, 101.75, 99.25, 101.5)],     # close = midpoint
    ]),
    # long white, three small blacks inside its range with falling closes,
    # long white opening above the last close and closing above 103
    "rise_fall3_methods": ([LONG_WHITE, (102.5, 102.75, 101.75, 102),
                             (102, 102.25, 101.25, 101.5), (101.5, 101.75, 100.75, 101),
                             (101.25, 104, 101, 103.75)], 100, [
        [LONG_WHITE, (102.5, 102.75, 101.75, 102), (102, 102.25, 101.25, 101.5),
         (101.5, 101.75, 100.75, 101), (101.25, 103.25, 101, 103)],    # close = close[i-4]
        [LONG_WHITE, (102.5, 102.75, 101.75, 102), (102, 102.25, 101.25, 102),
         (101.5, 101.75, 100.75, 101), (101.25, 104, 101, 103.75)],    # 2nd small close not lower
    ]),
    # black, white whose low is above the black close, black closing at the
    # first close within Equal
    "stick_sandwich": ([(102, 102.25, 99.75, 100), (100.5, 102.75, 100.25, 102.5),
                        (103, 103.25, 99.75, 100.0625)], 100, [
        [(102, 102.25, 99.75, 100), (100.5, 102.75, 100.25, 102.5), (103, 103.25, 99.75, 100.125)],
        [(102, 102.25, 99.75, 100), (100.5, 102.75, 100, 102.5), (103, 103.25, 99.75, 100.0625)],
    ]),
    # white, white gapping up, black opening inside the second body and
    # closing inside the gap, bodies within Near (0.38)
    "tasuki_gap": ([(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103),
                    (102.75, 103, 101.25, 101.5)], 100, [
        [(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103), (102, 102.25, 101, 101.25)],
    ]),
    # black, white engulfing it, higher close
    "three_outside": ([(101, 101.25, 99.75, 100), (99.75, 101.75, 99.5, 101.5),
                       (101.5, 102.25, 101.25, 102)], 100, [
        [(101, 101.25, 99.75, 100), (99.75, 101.75, 99.5, 101.5), (101.5, 102.25, 101.25, 101.5)],
        [(101, 101.25, 99.75, 100), (100, 101.75, 99.5, 101.5), (101.5, 102.25, 101.25, 102)],
    ]),
    # long white, black gapping up, black opening inside it and closing
    # inside the first body
    "two_crows": ([LONG_WHITE, (104.5, 104.75, 103.75, 104), (104.25, 104.5, 101.25, 101.5)],
                  -100, [
        [LONG_WHITE, (104.5, 104.75, 103.75, 104), (104.25, 104.5, 102.75, 103.0)],
        [LONG_WHITE, (104.5, 104.75, 103.75, 104), (104.5, 104.75, 101.25, 101.5)],  # open = open[i-1]
    ]),
    # long black, black with a higher close, lower low, open not above; small
    # white (< 1.25) opening above that low
    "unique_3_river": ([LONG_BLACK, (102, 102.25, 98.5, 100.5), (99, 99.75, 98.75, 99.5)],
                       100, [
        [LONG_BLACK, (102, 102.25, 98.5, 100.5), (98.5, 99.25, 98.5, 99.0)],  # open = low[i-1]
        [LONG_BLACK, (102, 102.25, 99.75, 100.5), (99, 99.75, 98.75, 99.5)],  # low not lower
    ]),
    # two whites with a body gap, black opening in the second and closing in
    # the first body
    "xside_gap_3_methods": ([(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103),
                             (102.5, 102.75, 100.25, 100.5)], 100, [
        [(100, 101.25, 99.75, 101), (102, 103.25, 101.75, 103), (102.5, 102.75, 99.75, 100.0)],
    ]),
    # inside bar, then a lower high and lower low: +100 at the pattern bar
    "hikkake": ([(100, 102, 98, 101), (100, 101, 99, 100.5), (99.5, 100.5, 98.5, 99)], 100, [
        [(100, 102, 98, 101), (100, 102, 99, 100.5), (99.5, 100.5, 98.5, 99)],  # no inside bar
    ]),
    # two nested inside bars, the middle one closing near its low (within
    # Near 0.64 of low 97), then a lower high and lower low
    "hikkake_mod": ([(100, 104, 96, 101), (101, 103, 97, 97.25), (99, 101, 98, 100),
                     (98.5, 100.5, 97.5, 98)], 100, [
        [(100, 104, 96, 101), (101, 103, 97, 98.0), (99, 101, 98, 100),
         (98.5, 100.5, 97.5, 98)],            # middle close 98 > 97.64
    ]),
}


def test_every_unpinned_pattern_has_a_fixture():
    pinned_elsewhere = {"three_white_soldiers", "three_black_crows", "identical_three_crows",
                        "kicking", "kicking_by_length", "three_stars_in_south",
                        "conceal_baby_swallow", "mat_hold", "stalled_pattern",
                        "three_line_strike", "doji", "marubozu", "spinning_top", "engulfing",
                        "long_line", "dragonfly_doji"}
    from tests.test_oracle_candles import names

    all_names = {n for n, _ in names()}
    assert set(CASES) == all_names - pinned_elsewhere
    assert len(CASES) == 45


@pytest.mark.parametrize("name", sorted(CASES))
def test_fires_with_stated_signal(name):
    candles, sig, _ = CASES[name]
    out = run(name, candles)
    assert out[-1] == sig, (name, int(out[-1]))


@pytest.mark.parametrize("name", sorted(CASES))
def test_near_miss_does_not_fire(name):
    _, _, misses = CASES[name]
    assert misses
    for k, candles in enumerate(misses):
        out = run(name, candles)
        assert out[-1] == 0, (name, k, int(out[-1]))


# ---- signals the single fixtures above do not reach ------------------------
def test_bearish_and_equal_end_variants():
    # black belt hold: upper shadow < 0.2
    assert run("belt_hold", [(103, 103, 99.75, 100)])[-1] == -100
    assert run("belt_hold", [(103, 103.25, 99.75, 100)])[-1] == 0
    # harami with one body end shared: +-80
    assert run("harami", [LONG_BLACK, (100, 101.25, 99.875, 101.0)])[-1] == 80
    assert run("harami", [LONG_WHITE, (101, 101.75, 100.25, 100.5)])[-1] == -100
    # bearish tristar: second doji gaps up, third's body top is lower
    assert run("tristar", [(100, 100.5, 99.5, 100.125), (101, 101.5, 100.5, 101.125),
                           (100.5, 101, 100, 100.625)])[-1] == -100
    # bearish three inside: long white, inside body, black below the first open
    assert run("three_inside", [LONG_WHITE, (101.5, 101.75, 100.75, 101.0),
                                (101, 101.25, 99.25, 99.5)])[-1] == -100
    # thrusting's midpoint is inclusive
    assert run("thrusting", [LONG_BLACK, (99.5, 101.75, 99.25, 101.5)])[-1] == -100
    # counterattack, bearish
    assert run("counterattack", [LONG_WHITE, (106, 106.25, 102.75, 103)])[-1] == -100


def test_hikkake_confirmations():
    base = [(100, 102, 98, 101), (100, 101, 99, 100.5), (99.5, 100.5, 98.5, 99)]
    # pattern +100, then a close above the inside bar's high (101) confirms
    out = run("hikkake", base + [(99.5, 101.75, 99.25, 101.5)])
    assert out[-2] == 100 and out[-1] == 200
    # a close at that high does not confirm
    out = run("hikkake", base + [(99.5, 101.25, 99.25, 101.0)])
    assert out[-1] == 0
    # confirmation on the third bar after the pattern still counts, the
    # fourth no longer
    wait = [(99.5, 100.25, 98.75, 100)] * 2
    out = run("hikkake", base + wait + [(99.5, 101.75, 99.25, 101.5)])
    assert out[-1] == 200
    out = run("hikkake", base + wait * 2 + [(99.5, 101.75, 99.25, 101.5)])
    assert out[-1] == 0
    mod = [(100, 104, 96, 101), (101, 103, 97, 97.25), (99, 101, 98, 100), (98.5, 100.5, 97.5, 98)]
    out = run("hikkake_mod", mod + [(99, 101.25, 98.75, 101.125)])
    assert out[-2] == 100 and out[-1] == 200
