"""GPU parity of ragged batches: every series of a ragged layout must equal
the oracle run on that series alone (window families bit-identical, scan
families within 1e-5 and at least 99.9% bit-identical)."""
import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _layout(lengths, seed=7):
    import tacuda_amd as ta

    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    d = ta.synth_ohlcv(int(max(lengths)), rows=len(lengths), seed=seed)
    cat = {k: np.concatenate([d[k][r, :n] for r, n in enumerate(lengths)]).astype(np.float32)
           for k in ("high", "low", "close")}
    return off, cat


def _rows(arr, off):
    return [arr[off[r]:off[r + 1]] for r in range(len(off) - 1)]


LENGTHS = {
    "short": [64, 100, 234, 256, 1, 0, 17, 255],
    "mixed": list(np.random.default_rng(1).integers(64, 2000, size=120)) + [0, 5000, 3],
    "long": [70001, 1025, 300000],
}

ONE = {
    "sma": ((20,), True), "ema": ((20,), False), "wma": ((20,), True), "dema": ((20,), False),
    "tema": ((20,), False), "rsi": ((14,), False), "stddev": ((20,), True),
    "max": ((14,), True),
}
HLC = {"atr": ((14,), False), "adx": ((14,), False), "willr": ((14,), True), "cci": ((14,), True)}


def _cmp(g, w, exact, name):
    if exact:
        compare(g, w, tol=0.0, min_exact=1.0, name=name)
    else:
        compare(g, w, tol=1e-5, min_exact=0.999, name=name)


def _cat_want(off, fn, *arrs):
    """Oracle per series, concatenated in layout order (empty series skipped)."""
    parts = []
    for r in range(len(off) - 1):
        sl = slice(off[r], off[r + 1])
        if off[r + 1] > off[r]:
            parts.append(fn(*[a[sl] for a in arrs]))
    if isinstance(parts[0], tuple):
        return tuple(np.concatenate([p[j] for p in parts]) for j in range(len(parts[0])))
    return np.concatenate(parts)


@pytest.mark.parametrize("case", sorted(LENGTHS))
def test_ragged_families(dev, case):
    from oracle import ORC
    from tacuda_amd.ops import Ragged

    print(f"ragged families {case}", flush=True)
    off, cat = _layout(LENGTHS[case])
    lay = Ragged(off)
    T = {k: torch.from_numpy(v).to(dev) for k, v in cat.items()}
    h, l, c = cat["high"], cat["low"], cat["close"]
    for name, (prm, exact) in ONE.items():
        got = lay.one(name, T["close"], *prm).cpu().numpy()
        want = _cat_want(off, lambda x: getattr(ORC, name)(x, *prm), c)
        _cmp(got, want, exact, name)
    for name, (prm, exact) in HLC.items():
        got = lay.hlc(name, T["high"], T["low"], T["close"], *prm).cpu().numpy()
        want = _cat_want(off, lambda a, b, d: getattr(ORC, name)(a, b, d, *prm), h, l, c)
        _cmp(got, want, exact, name)
    got = [t.cpu().numpy() for t in lay.macd(T["close"], 12, 26, 9)]
    want = _cat_want(off, lambda x: ORC.macd(x, 12, 26, 9), c)
    for j in range(3):
        _cmp(got[j], want[j], False, f"macd[{j}]")
    got = [t.cpu().numpy() for t in lay.stochastic(T["high"], T["low"], T["close"], 14, 3)]
    want = _cat_want(off, lambda a, b, d: ORC.stoch(a, b, d, 14, 3), h, l, c)
    for j in range(2):
        _cmp(got[j], want[j], True, f"stoch[{j}]")
    lay.close()


@pytest.mark.parametrize("case", sorted(LENGTHS))
def test_ragged_pack(dev, case):
    from oracle import ORC
    from tacuda_amd.ops import Ragged
    from tests.test_gpu_pack import DEFAULT, EXACT, _want

    print(f"ragged pack {case}", flush=True)
    off, cat = _layout(LENGTHS[case], seed=9)
    lay = Ragged(off)
    T = {k: torch.from_numpy(v).to(dev) for k, v in cat.items()}
    got = lay.pack(T["high"], T["low"], T["close"], **DEFAULT)
    torch.cuda.synchronize()
    for name, t in got.items():
        want = _cat_want(off, lambda a, b, d: _want(ORC, DEFAULT, a, b, d, name),
                         cat["high"], cat["low"], cat["close"])
        _cmp(t.cpu().numpy(), want, name in EXACT, name)
    lay.close()
