"""OBV and AD (running sums in double on the look-back scan engine) on the
GPU against the oracle's sequential loops (oracle/oracle.c orc_obv /
orc_ad).  The device sums associate differently from a sequential loop
(lane runs, then the look-back tree), as the reference's double scan does:
where every partial sum is exact (OBV with integer volumes) the outputs must
be bit-identical; otherwise at least 99.9% of the float outputs must be
bit-identical and all within 1e-6 relative of the running magnitude."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _close(got, want, scale, what):
    assert np.array_equal(np.isnan(got), np.isnan(want)), f"{what}: NaN masks differ"
    ok = ~np.isnan(want)
    same = np.mean(got[ok].view(np.uint32) == want[ok].view(np.uint32))
    assert same >= 0.999, f"{what}: only {same:.5f} bit-identical"
    err = np.abs(got[ok].astype(np.float64) - want[ok]) / np.maximum(scale[ok], 1e-30)
    assert err.max() <= 1e-6, f"{what}: max scaled error {err.max():.3g}"


def _bars(rows, n, seed, int_volume=False):
    import tacuda_amd as ta

    d = {k: np.asarray(v, np.float32).reshape(rows, n).copy()
         for k, v in ta.synth_ohlcv(n, rows=rows, seed=seed).items()}
    if int_volume:
        d["volume"] = np.round(d["volume"])
    d["close"][:, n // 5: n // 5 + 3] = d["close"][:, n // 5 - 1: n // 5]  # unchanged closes
    d["high"][:, n // 7] = d["low"][:, n // 7]  # a zero range
    return d


def _run(L, name, d, rows, n, dev, batch):
    keys = ("close", "volume") if name == "obv" else ("high", "low", "close", "volume")
    ins = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in keys]
    y = torch.empty((rows, n), device=dev)
    if batch:
        rc = getattr(L, f"ct_{name}_batch_d")(*[t.data_ptr() for t in ins], y.data_ptr(), n, rows,
                                             4, None)
    else:
        rc = getattr(L, f"ct_{name}_d")(*[t.data_ptr() for t in ins], y.data_ptr(), n, None)
    assert rc == 0, L.ct_get_last_error()
    torch.cuda.synchronize()
    return y.cpu().numpy(), keys


def _want(name, d, keys, r):
    from oracle import ORC

    return getattr(ORC, name)(*[d[k][r] for k in keys])


def _scale(name, d, keys, r):
    v = np.abs(d["volume"][r].astype(np.float64))
    # both sums are bounded by the running |volume| (NaN / Inf volumes aside)
    return np.cumsum(np.nan_to_num(v, nan=0.0, posinf=0.0))


@pytest.mark.parametrize("name", ["obv", "ad"])
@pytest.mark.parametrize("rows,n", [(1, 1_000_000), (1, 4099), (37, 5000), (300, 200)])
def test_against_oracle(dev, name, rows, n):
    import tacuda_amd as ta

    print(f"cum {name} {rows}x{n}", flush=True)
    L = ta.lib()
    d = _bars(rows, n, seed=rows + n)
    got, keys = _run(L, name, d, rows, n, dev, batch=rows > 1)
    for r in range(0, rows, max(1, rows // 25)):
        _close(got[r], _want(name, d, keys, r), _scale(name, d, keys, r), f"{name} row {r}")


@pytest.mark.parametrize("rows,n", [(1, 1_000_000), (50, 3000)])
def test_obv_exact_with_integer_volumes(dev, rows, n):
    import tacuda_amd as ta

    print(f"cum obv exact {rows}x{n}", flush=True)
    L = ta.lib()
    d = _bars(rows, n, seed=5, int_volume=True)
    got, keys = _run(L, "obv", d, rows, n, dev, batch=rows > 1)
    for r in range(0, rows, max(1, rows // 10)):
        want = _want("obv", d, keys, r)
        assert np.array_equal(got[r].view(np.uint32), want.view(np.uint32)), f"obv row {r}"


def test_nan_inputs_and_host_forms(dev):
    import tacuda_amd as ta

    print("cum NaN and host", flush=True)
    L = ta.lib()
    n = 3000
    d = _bars(1, n, seed=9)
    d["close"][0, 700] = np.nan  # OBV: no change counted at 700 and 701
    d["volume"][0, 1500] = np.nan  # OBV from 1500 on is NaN; AD skips the bar
    d["high"][0, 2000] = np.inf  # AD skips a non-finite bar
    for name in ("obv", "ad"):
        got, keys = _run(L, name, d, 1, n, dev, batch=False)
        want = _want(name, d, keys, 0)
        _close(got[0], want, _scale(name, d, keys, 0), f"{name} NaN")
        hy = np.empty(n, np.float32)
        rc = getattr(L, f"ct_{name}")(*[np.ascontiguousarray(d[k][0]).ctypes.data for k in keys],
                                      hy.ctypes.data, n, None)
        assert rc == 0
        assert np.array_equal(np.nan_to_num(hy, nan=-1.0), np.nan_to_num(got[0], nan=-1.0))
    assert L.ct_obv_d(None, None, None, 10, None) == 4
    x = torch.ones(10, device=dev)
    assert L.ct_ad_batch_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), 10,
                           0, 1, None) == 4
