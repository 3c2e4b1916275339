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


@pytest.mark.parametrize("case", sorted(LENGTHS))
def test_ragged_families(dev, case):
    from oracle import ORC
    from tacuda_amd.ops import Ragged

    print(f"ragged families {case}", flush=True)
    off, cat = _layout(LENGTHS[case])
    lay = Ragged(off)
    T = {k: torch.from_numpy(v).to(dev) for k, v in cat.items()}
    for name, (prm, exact) in ONE.items():
        got = lay.one(name, T["close"], *prm).cpu().numpy()
        for r, (g, x) in enumerate(zip(_rows(got, off), _rows(cat["close"], off))):
            if len(x):
                _cmp(g, getattr(ORC, name)(x, *prm), exact, f"{name} row {r}")
    for name, (prm, exact) in HLC.items():
        got = lay.hlc(name, T["high"], T["low"], T["close"], *prm).cpu().numpy()
        rows = zip(_rows(got, off), _rows(cat["high"], off), _rows(cat["low"], off),
                   _rows(cat["close"], off))
        for r, (g, h, l, c) in enumerate(rows):
            if len(c):
                _cmp(g, getattr(ORC, name)(h, l, c, *prm), exact, f"{name} row {r}")
    m, s, h = lay.macd(T["close"], 12, 26, 9)
    k, d = lay.stochastic(T["high"], T["low"], T["close"], 14, 3)
    for r, c in enumerate(_rows(cat["close"], off)):
        if not len(c):
            continue
        sl = slice(off[r], off[r + 1])
        want = ORC.macd(c, 12, 26, 9)
        for j, g in enumerate((m, s, h)):
            _cmp(g[sl].cpu().numpy(), want[j], False, f"macd[{j}] row {r}")
        want = ORC.stoch(cat["high"][sl], cat["low"][sl], c, 14, 3)
        _cmp(k[sl].cpu().numpy(), want[0], True, f"stoch k row {r}")
        _cmp(d[sl].cpu().numpy(), want[1], True, f"stoch d row {r}")
    lay.close()


@pytest.mark.parametrize("case", sorted(LENGTHS))
def test_ragged_pack(dev, case):
    from tacuda_amd.ops import Ragged
    from tests.test_gpu_pack import DEFAULT, EXACT, _want
    from oracle import ORC

    print(f"ragged pack {case}", flush=True)
    off, cat = _layout(LENGTHS[case], seed=9)
    lay = Ragged(off)
    T = {k: torch.from_numpy(v).to(dev) for k, v in cat.items()}
    got = lay.pack(T["high"], T["low"], T["close"], **DEFAULT)
    torch.cuda.synchronize()
    for name, t in got.items():
        g = t.cpu().numpy()
        for r in range(len(off) - 1):
            sl = slice(off[r], off[r + 1])
            if off[r + 1] == off[r]:
                continue
            w = _want(ORC, DEFAULT, cat["high"][sl], cat["low"][sl], cat["close"][sl], name)
            _cmp(g[sl], w, name in EXACT, f"{name} row {r}")
    lay.close()
