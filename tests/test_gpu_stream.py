"""Batch streams on the GPU: the HBM-resident state machines equal the host
ones bit for bit (host-array appends, device-pointer appends, and switching
execution mid-stream)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FAMS = [("ema", (20,)), ("rsi", (14,)), ("atr", (14,)), ("natr", (14,)), ("sma", (20,)),
        ("wma", (10,)), ("bbands", (20, 2.0, 1.5)), ("macd", (12, 26, 9)), ("obv", ()),
        ("ad", ())]


def _tup(x):
    return x if isinstance(x, tuple) else (x,)


def _eq(a, b, name):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), name
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), name


@pytest.mark.parametrize("family,prm", FAMS, ids=[f[0] for f in FAMS])
def test_gpu_batch_equals_cpu(dev, family, prm):
    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    print(f"stream batch {family}", flush=True)
    rows, n = 3000, 120
    d = ta.synth_ohlcv(n, rows=rows, seed=17)
    d["close"][7, 40] = np.nan
    cpu = StreamBatch(family, rows, *prm, exec="cpu")
    gpu = StreamBatch(family, rows, *prm, exec="gpu")
    for lo, hi in ((0, 1), (1, 60), (60, 61), (61, 120)):
        sl = (slice(None), slice(lo, hi))
        kw = dict(high=d["high"][sl], low=d["low"][sl], close=d["close"][sl],
                  volume=d["volume"][sl])
        for j, (a, b) in enumerate(zip(_tup(cpu.append(**kw)), _tup(gpu.append(**kw)))):
            _eq(b, a, f"{family} [{lo},{hi}) out {j}")
    assert gpu.bars == n


def test_device_append_and_exec_switch(dev):
    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    print("stream device append", flush=True)
    rows, n = 65536, 6
    d = ta.synth_ohlcv(n, rows=rows, seed=19)
    ref = StreamBatch("rsi", rows, 3, exec="cpu")
    want = ref.append(close=d["close"])
    b = StreamBatch("rsi", rows, 3, exec="auto")
    c = torch.from_numpy(np.ascontiguousarray(d["close"])).to(dev)
    got = []
    # bar 0 on the host (AUTO picks the GPU at this size), bars 1-3 on the
    # device, bar 4 forced to the CPU, bar 5 back on the GPU
    got.append(b.append(close=d["close"][:, 0:1]))
    for i in (1, 2, 3):
        col = c[:, i].contiguous()
        got.append(b.append_d(close=col).cpu().numpy()[:, None])
    b.set_exec("cpu")
    got.append(b.append(close=d["close"][:, 4:5]))
    b.set_exec("gpu")
    got.append(b.append(close=d["close"][:, 5:6]))
    _eq(np.concatenate(got, axis=1), want, "rsi mixed execution")
    assert b.bars == n
