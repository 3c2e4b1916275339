"""Writes candle_activation.json: the hand-authored OHLC bars of the
reference's candle activation fixture (benchmarks/differential_corpus.py:
73-142) -- a calm alternating series with injected sequences that must fire
the rare TA-Lib patterns -- and, for each injection, the pattern, the bar of
its last candle and the TA-Lib signal expected there.

Data only: the bar values are the fixture's documented numbers; nothing of
the reference is imported or run.  Run: python tests/golden/make_candle_fixture.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    n = 512
    o, h, l, c = [0.0] * n, [0.0] * n, [0.0] * n, [0.0] * n
    level = 100.0
    for i in range(n):
        d = 0.1 if i % 2 == 0 else -0.1
        o[i], c[i] = level - d, level + d
        h[i] = max(o[i], c[i]) + 0.15
        l[i] = min(o[i], c[i]) - 0.15

    def put(i0, bars):
        for k, (bo, bh, bl, bc) in enumerate(bars):
            o[i0 + k], h[i0 + k], l[i0 + k], c[i0 + k] = bo, bh, bl, bc

    put(40, [(100.00, 103.02, 99.90, 103.00), (101.00, 105.02, 100.90, 105.00),
             (103.00, 108.02, 102.90, 108.00)])
    put(80, [(100.00, 104.05, 99.95, 104.00), (103.50, 103.55, 100.49, 100.50),
             (102.50, 102.55, 98.99, 99.00), (100.00, 100.05, 96.99, 97.00)])
    put(120, [(104.00, 104.02, 99.99, 100.00), (100.00, 100.02, 95.99, 96.00),
              (96.00, 96.02, 91.99, 92.00)])
    put(160, [(100.000, 100.005, 95.995, 96.000), (101.000, 105.005, 100.995, 105.000)])
    put(200, [(100.00, 100.02, 91.00, 96.00), (98.00, 98.02, 93.00, 95.50),
              (95.40, 95.42, 94.78, 94.80)])
    put(240, [(100.000, 100.005, 95.995, 96.000), (95.500, 95.505, 91.495, 91.500),
              (90.50, 92.50, 88.99, 89.00), (93.00, 93.01, 88.49, 88.50)])
    put(280, [(100.00, 105.01, 99.99, 105.00), (105.90, 106.00, 105.75, 105.80),
              (105.05, 105.15, 104.90, 104.95), (104.80, 104.90, 104.60, 104.70),
              (105.00, 108.01, 104.99, 108.00)])
    put(320, [(100.00, 103.01, 99.99, 103.00), (101.50, 106.01, 101.40, 106.00),
              (105.90, 106.45, 105.85, 106.40)])
    put(360, [(100.00, 103.02, 99.90, 103.00), (101.00, 105.02, 100.90, 105.00),
              (103.00, 108.02, 102.90, 108.00), (108.50, 108.60, 99.40, 99.50)])
    # (pattern, last bar of the injection, TA-Lib signal there)
    fires = [
        ("three_white_soldiers", 42, 100),
        ("three_black_crows", 83, -100),
        ("identical_three_crows", 122, -100),
        ("kicking", 161, 100),
        ("kicking_by_length", 161, -100),  # equal bodies: the first candle's color
        ("three_stars_in_south", 202, 100),
        ("conceal_baby_swallow", 243, 100),
        ("mat_hold", 284, 100),
        ("stalled_pattern", 322, -100),
        ("three_line_strike", 363, 100),
    ]
    return {"open": o, "high": h, "low": l, "close": c, "fires": fires}


if __name__ == "__main__":
    with open(os.path.join(HERE, "candle_activation.json"), "w") as f:
        json.dump(build(), f)
