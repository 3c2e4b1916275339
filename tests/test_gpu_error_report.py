"""Per-output error of the HIP path against the oracle on the benchmark
workloads' inputs (the seeded generator at the benchmark's shapes, sampled
rows where the oracle would take minutes): NaN masks identical, max relative
error |a-b|/max(1,|a|,|b|), max absolute error and the fraction of outputs
bit-identical to the oracle's float32 result.  Every output must meet the
parity bar (1e-5 relative, 99.9% bit-identical; candles and window families
bit-identical); the table is printed and, when TACUDA_PARITY_REPORT names a
file, written there as JSON."""
import json
import os

import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REPORT = {}


def _stats(label, got, want, exact_only=False):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    st = compare(got, want, tol=0.0 if exact_only else 1e-5,
                 min_exact=1.0 if exact_only else 0.999, name=label)
    ok = ~np.isnan(want) & np.isfinite(want)
    abs_err = float(np.max(np.abs(got[ok].astype(np.float64) - want[ok]))) if ok.any() else 0.0
    REPORT[label] = {"n": st["n"], "bit_identical": st["exact"], "max_rel": st["max_rel"],
                     "max_abs": abs_err}


@pytest.fixture(scope="module", autouse=True)
def _write_report():
    yield
    lines = [f"{k:<48} n={v['n']:>10} exact={v['bit_identical']:.6f} "
             f"max_rel={v['max_rel']:.3g} max_abs={v['max_abs']:.3g}" for k, v in REPORT.items()]
    print("\n" + "\n".join(lines), flush=True)
    path = os.environ.get("TACUDA_PARITY_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump(REPORT, f, indent=1, sort_keys=True)


def test_pipeline7_one_million_bars(dev):
    """The 7-indicator pipeline on 10^6 bars, fused pass and drop-in calls."""
    import tacuda_amd as ta
    from oracle import ORC

    print("report pipeline7", flush=True)
    d = ta.synth_ohlcv(1_000_000, seed=0)
    h, l, c = (torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close"))
    p = dict(sma_period=20, ema_period=20, rsi_period=14, macd_fast=12, macd_slow=26,
             macd_signal=9, bb_period=20, bb_up=2.0, bb_dn=2.0, atr_period=14, stoch_k=14,
             stoch_d=3)
    outs = ("sma", "ema", "rsi", "macd", "macd_signal", "macd_hist", "bb_upper", "bb_middle",
            "bb_lower", "atr", "stoch_k", "stoch_d")
    fused = ta.ops.pack(h, l, c, outputs=outs, **p)
    drop = {"sma": ta.ops.sma(c, 20), "ema": ta.ops.ema(c, 20), "rsi": ta.ops.rsi(c, 14),
            "atr": ta.ops.atr(h, l, c, 14)}
    m = ta.ops.macd(c, 12, 26, 9)
    drop.update(macd=m[0], macd_signal=m[1], macd_hist=m[2])
    b = ta.ops.bbands(c, 20, 2.0, 2.0)
    drop.update(bb_upper=b[0], bb_middle=b[1], bb_lower=b[2])
    s = ta.ops.stochastic(h, l, c, 14, 3)
    drop.update(stoch_k=s[0], stoch_d=s[1])
    torch.cuda.synchronize()
    H, L, C = d["high"], d["low"], d["close"]
    mw, bw, sw = ORC.macd(C, 12, 26, 9), ORC.bbands(C, 20, 2.0, 2.0), ORC.stoch(H, L, C, 14, 3)
    want = {"sma": ORC.sma(C, 20), "ema": ORC.ema(C, 20), "rsi": ORC.rsi(C, 14),
            "macd": mw[0], "macd_signal": mw[1], "macd_hist": mw[2], "bb_upper": bw[0],
            "bb_middle": bw[1], "bb_lower": bw[2], "atr": ORC.atr(H, L, C, 14),
            "stoch_k": sw[0], "stoch_d": sw[1]}
    exact = {"sma", "bb_upper", "bb_middle", "bb_lower", "stoch_k", "stoch_d"}
    for k in outs:
        _stats(f"pipeline7 fused {k}", fused[k].cpu().numpy(), want[k], exact_only=k in exact)
        _stats(f"pipeline7 drop-in {k}", drop[k].cpu().numpy(), want[k])


def test_batch13(dev):
    """The 13-indicator fused batch, 4000 x 234."""
    import tacuda_amd as ta
    from oracle import ORC
    from tests.test_gpu_pack import DEFAULT, EXACT, _want

    print("report batch13", flush=True)
    d = ta.synth_ohlcv(234, rows=4000, seed=0)
    T = [torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close")]
    got = ta.ops.pack(*T, **DEFAULT)
    torch.cuda.synchronize()
    H, L, C = d["high"], d["low"], d["close"]
    for name, t in got.items():
        w = np.stack([_want(ORC, DEFAULT, H[r], L[r], C[r], name) for r in range(4000)])
        _stats(f"batch13 {name}", t.cpu().numpy(), w, exact_only=name in EXACT)


def test_scan256_sampled_rows(dev):
    """The fused long scan on 256 x 2^20, eight rows checked."""
    import tacuda_amd as ta
    from oracle import ORC
    from tests.test_gpu_scan import DEFAULT, _want

    print("report scan256", flush=True)
    c = ta.synth_ohlcv(1 << 20, rows=256, seed=0)["close"]
    got = ta.ops.scan(torch.from_numpy(c).to(dev))
    torch.cuda.synchronize()
    rows = [0, 37, 101, 128, 170, 199, 230, 255]
    for name, t in got.items():
        g = t[rows].cpu().numpy()
        w = np.stack([_want(ORC, DEFAULT, c[r], name) for r in rows])
        _stats(f"scan256 {name} (8 rows)", g, w)


def test_candles_sampled_rows(dev):
    """All 61 patterns, int32, 1024 x 16,384: 16 rows checked bit for bit."""
    import tacuda_amd as ta
    from tests.test_gpu_candles import _oracle_all

    print("report candles", flush=True)
    d = ta.synth_ohlcv(16384, rows=1024, seed=0)
    x = [torch.from_numpy(d[k]).to(dev) for k in ("open", "high", "low", "close")]
    got = ta.ops.cdl_all(*x)
    torch.cuda.synchronize()
    rows = list(range(0, 1024, 64))
    g = got.cpu().numpy()
    mism = 0
    total = 0
    for r in rows:
        w = _oracle_all(d["open"][r], d["high"][r], d["low"][r], d["close"][r])
        gr = g[:, r, :] if g.ndim == 3 else g[r]
        mism += int(np.sum(gr != w))
        total += w.size
    assert mism == 0, f"{mism} candle signals differ"
    REPORT["candles 61 x 16 rows"] = {"n": total, "bit_identical": 1.0, "max_rel": 0.0,
                                      "max_abs": 0.0}
