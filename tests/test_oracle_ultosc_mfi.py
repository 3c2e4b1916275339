"""Pins of the oracle's ULTOSC and MFI (oracle/oracle.c): exact values on
integer bars (where the float sums and quotients are the formula's), numpy
brute force, the period sorting, and monotone / degenerate closed forms."""
import numpy as np
import pytest

from oracle import ORC


def _bars(n, seed):
    import tacuda_amd as ta

    return {k: np.asarray(v, np.float32) for k, v in ta.synth_ohlcv(n, seed=seed).items()}


def _ultosc_numpy(h, l, c, p1, p2, p3):
    """Independent restatement in float64 (then float32): buying pressure and
    true range against the previous close, averaged over three windows."""
    n = len(c)
    pc = np.concatenate([c[:1], c[:-1]]).astype(np.float64)
    lo = np.minimum(l.astype(np.float64), pc)
    bp = c - lo
    tr = np.maximum(h.astype(np.float64), pc) - lo
    out = np.full(n, np.nan)
    for t in range(p3, n):
        a = []
        for p in (p1, p2, p3):
            b, r = bp[t - p + 1:t + 1].sum(), tr[t - p + 1:t + 1].sum()
            a.append(0.0 if r == 0 else b / r)
        out[t] = 100.0 * (4 * a[0] + 2 * a[1] + a[2]) / 7.0
    return out


def test_ultosc_integer_bars_exact():
    rng = np.random.default_rng(1)
    n = 400
    lo = rng.integers(100, 200, n).astype(np.float32)
    hi = lo + rng.integers(0, 8, n).astype(np.float32)
    cl = lo + np.floor(rng.random(n) * (hi - lo + 1)).astype(np.float32)
    got = ORC.ultosc(hi, lo, cl, 7, 14, 28)
    want = _ultosc_numpy(hi, lo, cl, 7, 14, 28)
    assert np.isnan(got[:28]).all()
    # sums of small integers are exact in float; the quotients and the final
    # combination round once or twice in float: within 2 float ulps
    np.testing.assert_allclose(got[28:], want[28:], rtol=3e-7, atol=2e-5)


def test_ultosc_brute_force_and_sorting():
    d = _bars(1500, 2)
    h, l, c = d["high"], d["low"], d["close"]
    got = ORC.ultosc(h, l, c, 7, 14, 28)
    np.testing.assert_allclose(got[28:], _ultosc_numpy(h, l, c, 7, 14, 28)[28:], rtol=0,
                               atol=2e-3)
    # periods arrive in any order
    assert np.array_equal(np.nan_to_num(ORC.ultosc(h, l, c, 28, 7, 14)), np.nan_to_num(got))
    # closes at the high of bars that gap nowhere: pressure = range, ULTOSC 100
    x = np.linspace(100, 120, 200).astype(np.float32)
    assert np.all(ORC.ultosc(x, x - 1, x, 2, 3, 5)[5:] == 100.0)
    assert np.isnan(ORC.ultosc(h[:28], l[:28], c[:28], 7, 14, 28)).all()  # p3 >= size


def _mfi_numpy(h, l, c, v, p):
    tp = (h.astype(np.float64) + l + c) / 3.0
    out = np.full(len(c), np.nan)
    for t in range(p, len(c)):
        pos = neg = 0.0
        for k in range(t - p + 1, t + 1):
            f = tp[k] * v[k]
            if tp[k] > tp[k - 1]:
                pos += f
            elif tp[k] < tp[k - 1]:
                neg += f
        tot = pos + neg
        out[t] = 0.0 if tot < 1.0 else 100.0 * pos / tot
    return out


@pytest.mark.parametrize("p", [2, 14, 40])
def test_mfi_brute_force(p):
    d = _bars(800, 3)
    got = ORC.mfi(d["high"], d["low"], d["close"], d["volume"], p)
    want = _mfi_numpy(d["high"], d["low"], d["close"], d["volume"], p)
    assert np.isnan(got[:p]).all()
    np.testing.assert_allclose(got[p:], want[p:].astype(np.float32), rtol=1e-6, atol=1e-5)


def test_mfi_closed_forms():
    x = np.linspace(50, 80, 100).astype(np.float32)
    v = np.full(100, 1000.0, np.float32)
    assert np.all(ORC.mfi(x, x, x, v, 10)[10:] == 100.0)  # typical price always rising
    assert np.all(ORC.mfi(x[::-1].copy(), x[::-1].copy(), x[::-1].copy(), v, 10)[10:] == 0.0)
    flat = np.full(100, 5.0, np.float32)
    assert np.all(ORC.mfi(flat, flat, flat, v, 10)[10:] == 0.0)  # no flow: total < 1
    with pytest.raises(ValueError):
        ORC.mfi(x, x, x, v, 1)
