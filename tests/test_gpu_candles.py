"""GPU parity of the 61 candlestick patterns: int32 signals bit-exact
against the oracle for every pattern, through the all-61 launch, the
single-pattern int32 call and the reference's float ct_cdl_* calls; tile
boundaries, batches, NaN holes and the activation fixture."""
import numpy as np
import pytest

from tests.candle_data import activation, quantized, with_holes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oracle_all(o, h, l, c):
    from oracle import ORC
    from tacuda_amd.ops import cdl_names

    names = cdl_names()
    if c.ndim == 1:
        return np.stack([ORC.candle(n, o, h, l, c) for n in names])
    return np.stack([np.stack([ORC.candle(n, o[r], h[r], l[r], c[r]) for r in range(c.shape[0])])
                     for n in names])


def _dev(arrs, dev):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrs]


def _check_all(arrs, dev, tag):
    from tacuda_amd import ops

    got = ops.cdl_all(*_dev(arrs, dev)).cpu().numpy()
    want = _oracle_all(*arrs)
    names = ops.cdl_names()
    for k in range(61):
        bad = np.flatnonzero(got[k].ravel() != want[k].ravel())
        assert bad.size == 0, f"{tag} {names[k]}: {bad.size} mismatches, first at {bad[:5]}"
    return got


@pytest.mark.parametrize("n", [1, 14, 255, 256, 257, 1000, 5000])
def test_all_patterns_single_series(dev, n):
    print(f"candles single n={n}", flush=True)
    _check_all(quantized(n, seed=7), dev, f"n={n}")


@pytest.mark.parametrize("shape", [(16, 3000), (300, 234)])
def test_all_patterns_batch(dev, shape):
    rows, n = shape
    print(f"candles batch {rows}x{n}", flush=True)
    _check_all(quantized(n, rows=rows, seed=9), dev, f"{rows}x{n}")


def test_nan_holes(dev):
    print("candles nan holes", flush=True)
    _check_all(with_holes(quantized(3000, seed=11)), dev, "holes")


def test_activation_fixture_on_gpu(dev):
    from tacuda_amd import ops

    print("candles activation", flush=True)
    arrs, fires = activation()
    got = _check_all(arrs, dev, "activation")
    names = ops.cdl_names()
    for name, bar, sig in fires:
        assert got[names.index(name)][bar] == sig, name


@pytest.mark.parametrize("name", ["engulfing", "hikkake", "hikkake_mod", "mat_hold", "doji",
                                  "three_line_strike", "harami"])
def test_single_pattern_and_legacy_float_calls(dev, name):
    from oracle import ORC
    from tacuda_amd import ops

    print(f"candles legacy {name}", flush=True)
    o, h, l, c = quantized(4000, seed=13)
    want = ORC.candle(name, o, h, l, c)
    T = _dev((o, h, l, c), dev)
    assert np.array_equal(ops.cdl(name, *T).cpu().numpy(), want)
    assert np.array_equal(ops.cdl_legacy(name, *T).cpu().numpy(), want.astype(np.float32))
    ob, hb, lb, cb = quantized(700, rows=5, seed=15)
    want = np.stack([ORC.candle(name, ob[r], hb[r], lb[r], cb[r]) for r in range(5)])
    got = ops.cdl_legacy(name, *_dev((ob, hb, lb, cb), dev)).cpu().numpy()
    assert np.array_equal(got, want.astype(np.float32))
