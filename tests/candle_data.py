"""Candle test inputs: the activation fixture, tick-quantized random walks
(prices on a 0.05 grid, so equal opens/closes and exact body matches occur)
and NaN holes."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def activation():
    d = json.load(open(os.path.join(HERE, "golden", "candle_activation.json")))
    arrs = [np.array(d[k], np.float32) for k in ("open", "high", "low", "close")]
    return arrs, [tuple(f) for f in d["fires"]]


def quantized(n, rows=1, seed=0, tick=0.05):
    """Seeded OHLCV with opening gaps (open = previous close * (1 + N(0,
    0.004))), on a price grid."""
    import tacuda_amd as ta

    d = ta.synth_ohlcv(n, rows=rows, seed=seed)
    rng = np.random.default_rng(seed + 100)
    gap = 1.0 + 0.004 * rng.standard_normal(d["open"].shape)
    q = lambda x: (np.round(x.astype(np.float64) / tick) * tick).astype(np.float32)  # noqa: E731
    o, h, l, c = q(d["open"] * gap), q(d["high"]), q(d["low"]), q(d["close"])
    h = np.maximum(h, np.maximum(o, c))
    l = np.minimum(l, np.minimum(o, c))
    return o, h, l, c


def with_holes(arrs, every=97):
    out = [a.copy() for a in arrs]
    for a in out:
        a[..., ::every] = np.nan
    return out
