"""GPU checks of the kernel timer (ctx_kernel_timing / ctx_kernel_times):
the fused pass, the long scan and the candle launch name their kernels and
time them with events on the caller's stream; nothing is recorded while
timing is off or while the stream is captured."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _pipeline(dev):
    import tacuda_amd as ta

    d = ta.synth_ohlcv(20000, seed=3)
    return [torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close")]


def test_two_pass_kernels_are_timed(dev):
    import tacuda_amd as ta

    print("timing two-pass", flush=True)
    h, l, c = _pipeline(dev)
    outs = ("sma", "ema", "rsi", "macd", "atr", "stoch_k")
    ta.ops.pack(h, l, c, outputs=outs)
    torch.cuda.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        ta.ops.pack(h, l, c, outputs=outs)
        ta.ops.pack(h, l, c, outputs=outs)
    finally:
        ta.kernel_timing(False)
    rec = ta.kernel_times()
    names = [k for k, _ in rec]
    # one row of 40 segments: the emit waves scan the row's group composites
    # themselves, so there is no k_pipe_rows launch
    one = ["k_pipe_fold", "k_pipe_group", "k_pipe_emit"]
    assert names == one + one
    assert all(ms > 0.0 for _, ms in rec)
    ta.ops.pack(h, l, c, outputs=outs)
    torch.cuda.synchronize()
    assert ta.kernel_times() == []


def test_one_launch_pass_is_timed(dev):
    import tacuda_amd as ta

    print("timing one-launch pass", flush=True)
    h, l, c = _pipeline(dev)
    ta.ops.rsi(c, 14)
    torch.cuda.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        ta.ops.pack(h, l, c, outputs=("rsi",))
        ta.ops.pack(h, l, c, outputs=("macd", "macd_hist"))
    finally:
        ta.kernel_timing(False)
    rec = ta.kernel_times()
    # one recurrence group and no window part: one launch each
    assert [k for k, _ in rec] == ["k_pipe_solo", "k_pipe_solo"]
    assert all(ms > 0.0 for _, ms in rec)


def test_scan_and_candles_are_timed(dev):
    import tacuda_amd as ta

    print("timing scan, candles", flush=True)
    d = ta.synth_ohlcv(1 << 14, rows=16, seed=5)
    c = torch.from_numpy(d["close"]).to(dev)
    x = [torch.from_numpy(d[k]).to(dev) for k in ("open", "high", "low", "close")]
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        ta.ops.scan(c)
        ta.ops.cdl_all(*x)
    finally:
        ta.kernel_timing(False)
    names = [k for k, _ in ta.kernel_times()]
    assert names[:2] == ["k_scan_lead", "k_scan_one"]
    assert names[2:] == ["k_cdl<I32,ALL>"]


def test_captured_launches_are_not_timed(dev):
    import tacuda_amd as ta

    print("timing capture", flush=True)
    h, l, c = _pipeline(dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ref = ta.ops.pack(h, l, c, outputs=("ema", "sma"))
    s.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            out = ta.ops.pack(h, l, c, outputs=("ema", "sma"))
        g.replay()
        torch.cuda.synchronize()
    finally:
        ta.kernel_timing(False)
    assert ta.kernel_times() == []
    for k in ref:
        assert np.array_equal(ref[k].cpu().numpy().view(np.uint32),
                              out[k].cpu().numpy().view(np.uint32))
