"""MACDEXT (a moving-average type per line, the reference's formula-backend
path) and TRIMA on the GPU against the oracle, which is itself pinned to
the vendored TA-Lib 0.8.3 (TA_MACDEXT, TA_TRIMA): single series, dense
batches, host calls, ct_macd with non-EMA types, leading NaN prefixes."""
import ctypes

import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TYPES = [0, 1, 2, 3, 4, 5, 6, 8]  # SMA EMA WMA DEMA TEMA TRIMA KAMA T3


def _near(g, w, name):
    # both sides evaluate TA-Lib's arithmetic in double with the same order
    compare(g, w, tol=1e-7, min_exact=0.999, name=name)


@pytest.mark.parametrize("sgt", TYPES)
def test_macdext_every_type_combination(dev, sgt):
    import tacuda_amd as ta
    from oracle import ORC

    print(f"macdext signal type {sgt}", flush=True)
    d = ta.synth_ohlcv(3000, seed=31)
    c = d["close"]
    x = torch.from_numpy(c).to(dev)
    for ft in TYPES:
        for st in TYPES:
            got = ta.ops.macdext(x, 12, ft, 26, st, 9, sgt)
            want = ORC.macdext(c, 12, ft, 26, st, 9, sgt)
            for g, w, k in zip(got, want, ("macd", "signal", "hist")):
                _near(g.cpu().numpy(), w, f"macdext {ft}/{st}/{sgt} {k}")


@pytest.mark.parametrize("args", [(26, 3, 12, 0, 9, 4), (5, 8, 3, 6, 1, 5), (2, 2, 30, 5, 7, 8),
                                  (12, 1, 26, 1, 9, 1)])
def test_macdext_batch_prefix_and_host(dev, args):
    """Reversed periods, one-bar signal, all-EMA; a dense batch whose rows
    carry different NaN prefixes; the host forms."""
    import tacuda_amd as ta
    from oracle import ORC

    print(f"macdext batch {args}", flush=True)
    L = ta.lib()
    rows, n = 24, 900
    d = ta.synth_ohlcv(n, rows=rows, seed=32)
    c = np.ascontiguousarray(d["close"])
    for r in range(rows):
        c[r, : (r * 37) % 300] = np.nan
    X = torch.from_numpy(c).to(dev)
    got = [t.cpu().numpy() for t in ta.ops.macdext(X, *args)]
    for r in range(rows):
        want = ORC.macdext(c[r], *args)
        for g, w, k in zip(got, want, ("macd", "signal", "hist")):
            _near(g[r], w, f"macdext {args} row {r} {k}")
    hm, hs, hh = (np.empty_like(c) for _ in range(3))
    p = lambda a: a.ctypes.data  # noqa: E731
    assert L.ct_macdext_batch(p(c), p(hm), p(hs), p(hh), n, rows, *args, 0, None) == 0
    for g, h in zip(got, (hm, hs, hh)):
        assert np.array_equal(g.view(np.uint32), h.view(np.uint32))
    one = np.ascontiguousarray(c[5])
    om, os_, oh = (np.empty_like(one) for _ in range(3))
    assert L.ct_macdext(p(one), p(om), p(os_), p(oh), n, *args, None) == 0
    for g, h in zip(got, (om, os_, oh)):
        assert np.array_equal(g[5].view(np.uint32), h.view(np.uint32))


@pytest.mark.parametrize("ma_type", [0, 2, 3, 4, 5, 6, 8])
def test_macd_with_non_ema_type(dev, ma_type):
    """ct_macd_d with a moving-average type other than EMA is MACDEXT with
    that type on all three lines (reference MACDEXT.cu:164-235)."""
    import tacuda_amd as ta
    from oracle import ORC

    print(f"macd type {ma_type}", flush=True)
    c = ta.synth_ohlcv(5000, seed=33)["close"]
    x = torch.from_numpy(c).to(dev)
    got = ta.ops.macd(x, 12, 26, 9, ma_type)
    want = ORC.macdext(c, 12, ma_type, 26, ma_type, 9, ma_type)
    for g, w, k in zip(got, want, ("macd", "signal", "hist")):
        _near(g.cpu().numpy(), w, f"macd type {ma_type} {k}")


@pytest.mark.parametrize("p", [1, 2, 3, 10, 11, 30, 200])
def test_trima_bit_identical(dev, p):
    import tacuda_amd as ta
    from oracle import ORC

    print(f"trima {p}", flush=True)
    L = ta.lib()
    d = ta.synth_ohlcv(20000, seed=34)
    c = d["close"].copy()
    c[5000] = np.nan
    x = torch.from_numpy(c).to(dev)
    compare(ta.ops.trima(x, p).cpu().numpy(), ORC.trima(c, p), tol=0.0, min_exact=1.0,
            name=f"trima {p}")
    b = ta.synth_ohlcv(700, rows=9, seed=35)["close"]
    B = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    got = ta.ops.trima(B, p).cpu().numpy()
    for r in range(9):
        compare(got[r], ORC.trima(b[r], p), tol=0.0, min_exact=1.0, name=f"trima {p} row {r}")
    hy = np.empty_like(c)
    assert L.ct_trima(c.ctypes.data, hy.ctypes.data, len(c), p, None) == 0
    compare(hy, ORC.trima(c, p), tol=0.0, min_exact=1.0, name=f"host trima {p}")
    y = torch.empty_like(x)
    assert L.ct_ma_device(x.data_ptr(), y.data_ptr(), len(c), p, 5,
                          torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    compare(y.cpu().numpy(), ORC.trima(c, p), tol=0.0, min_exact=1.0, name=f"ma trima {p}")
