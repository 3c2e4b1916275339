"""The allocation-free fused pass (ctx_pack_batch_noalloc_d) through the C
ABI: two calls in a row, the second with a smaller input, each bit-identical
to ctx_pack_batch_d on the same inputs, for the single-launch pass (4000 x
234, all 18 outputs) and the two-pass long-row pass (one row of 200,000
bars, then 50,000); the call allocates nothing (the workspace cache's
allocation count does not move); a pass larger than its scratch is refused
with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bind(L):
    import tacuda_amd as ta

    fn = L.ctx_pack_batch_noalloc_d
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 2 + [
        ctypes.POINTER(ta.PackParams), ctypes.POINTER(ta.PackOutputs), ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def _run(fn_name, T, n, rows, params, outputs, dev):
    import tacuda_amd as ta

    L = ta.lib()
    fn = _bind(L) if fn_name == "noalloc" else L.ctx_pack_batch_d
    res = {k: torch.full((rows, n), 7.0, device=dev) for k in outputs}
    outs = ta.PackOutputs(**{k: v.data_ptr() for k, v in res.items()})
    rc = fn(*[t.data_ptr() for t in T], n, rows, params, outs,
            torch.cuda.current_stream().cuda_stream)
    assert rc == 0, (fn_name, rc, L.ct_get_last_error())
    torch.cuda.synchronize()
    return res


def _same(a, b, name):
    for k in a:
        assert torch.equal(a[k].view(torch.int32), b[k].view(torch.int32)), f"{name} {k}"


@pytest.mark.parametrize("shapes", [((4000, 234), (300, 100)), ((1, 200000), (1, 50000))],
                         ids=["batch", "long-row"])
def test_noalloc_equals_pack_on_two_calls(dev, shapes):
    import tacuda_amd as ta
    from tacuda_amd import PACK_OUTPUTS
    from tacuda_amd.ops import PACK_DEFAULTS

    L = ta.lib()
    prm = ta.PackParams(**PACK_DEFAULTS)
    for i, (rows, n) in enumerate(shapes):
        print(f"noalloc call {i}: {rows}x{n}", flush=True)
        d = ta.synth_ohlcv(n, rows=rows, seed=60 + i)
        T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev).reshape(rows, n)
             for k in ("high", "low", "close")]
        outputs = PACK_OUTPUTS if rows > 1 else ("sma", "ema", "rsi", "macd", "macd_signal",
                                                  "macd_hist", "bb_upper", "bb_middle",
                                                  "bb_lower", "atr", "stoch_k", "stoch_d")
        before = L.ct_device_buffer_pool_allocation_count()
        got = _run("noalloc", T, n, rows, prm, outputs, dev)
        assert L.ct_device_buffer_pool_allocation_count() == before
        want = _run("pack", T, n, rows, prm, outputs, dev)
        _same(got, want, f"{rows}x{n}")


def test_noalloc_refuses_a_pass_beyond_its_scratch(dev):
    import tacuda_amd as ta
    from tacuda_amd.ops import PACK_DEFAULTS

    print("noalloc oversize", flush=True)
    L = ta.lib()
    fn = _bind(L)
    n = 40_000_000
    x = torch.ones(n, device=dev)
    y = torch.empty(n, device=dev)
    prm = ta.PackParams(**PACK_DEFAULTS)
    outs = ta.PackOutputs(rsi=y.data_ptr())
    rc = fn(x.data_ptr(), x.data_ptr(), x.data_ptr(), n, 1, prm, outs,
            torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 4
    assert b"scratch" in L.ct_get_last_error()
