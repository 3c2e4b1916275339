"""ULTOSC and MFI on the GPU against the oracle: device single calls, dense
batches, host forms, period edges and errors.  ULTOSC is float arithmetic in
the reference's order and must match bit for bit; MFI sums double flows per
window in the oracle's order and must match bit for bit too."""
import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bars(rows, n, seed):
    import tacuda_amd as ta

    d = {k: np.asarray(v, np.float32).reshape(rows, n).copy()
         for k, v in ta.synth_ohlcv(n, rows=rows, seed=seed).items()}
    d["close"][:, n // 3] = np.nan
    d["volume"][:, n // 4] = 0.0
    return d


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("rows,n", [(1, 100003), (23, 777)])
def test_ultosc_and_mfi(dev, rows, n):
    import tacuda_amd as ta
    from oracle import ORC

    L = ta.lib()
    d = _bars(rows, n, seed=rows)
    h, l, c, v = (_dev(d[k], dev) for k in ("high", "low", "close", "volume"))
    for (p1, p2, p3) in ((7, 14, 28), (20, 3, 9)):
        print(f"ultosc {rows}x{n} {p1},{p2},{p3}", flush=True)
        y = torch.empty((rows, n), device=dev)
        if rows == 1:
            rc = L.ct_ultosc_d(h.data_ptr(), l.data_ptr(), c.data_ptr(), y.data_ptr(), n, p1, p2,
                               p3, None)
        else:
            rc = L.ct_ultosc_batch_d(h.data_ptr(), l.data_ptr(), c.data_ptr(), y.data_ptr(), n,
                                     rows, p1, p2, p3, 4, None)
        assert rc == 0, L.ct_get_last_error()
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        for r in range(0, rows, max(1, rows // 8)):
            want = ORC.ultosc(d["high"][r], d["low"][r], d["close"][r], p1, p2, p3)
            compare(got[r], want, tol=0.0, min_exact=1.0, name=f"ultosc row {r}")
    for p in (2, 14):
        print(f"mfi {rows}x{n} {p}", flush=True)
        y = torch.empty((rows, n), device=dev)
        if rows == 1:
            rc = L.ct_mfi_d(h.data_ptr(), l.data_ptr(), c.data_ptr(), v.data_ptr(), y.data_ptr(), n,
                            p, None)
        else:
            rc = L.ct_mfi_batch_d(h.data_ptr(), l.data_ptr(), c.data_ptr(), v.data_ptr(),
                                  y.data_ptr(), n, rows, p, 4, None)
        assert rc == 0, L.ct_get_last_error()
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        for r in range(0, rows, max(1, rows // 8)):
            want = ORC.mfi(d["high"][r], d["low"][r], d["close"][r], d["volume"][r], p)
            compare(got[r], want, tol=0.0, min_exact=1.0, name=f"mfi row {r}")


def test_host_forms_and_edges(dev):
    import tacuda_amd as ta
    from oracle import ORC

    print("ultosc/mfi host and edges", flush=True)
    L = ta.lib()
    n = 2000
    d = _bars(1, n, seed=4)
    hh, hl, hc, hv = (np.ascontiguousarray(d[k][0]) for k in ("high", "low", "close", "volume"))
    y = np.empty(n, np.float32)
    assert L.ct_ultosc(hh.ctypes.data, hl.ctypes.data, hc.ctypes.data, y.ctypes.data, n, 5, 10,
                       20, None) == 0
    compare(y, ORC.ultosc(hh, hl, hc, 5, 10, 20), tol=0.0, min_exact=1.0, name="ultosc host")
    assert L.ct_mfi(hh.ctypes.data, hl.ctypes.data, hc.ctypes.data, hv.ctypes.data, y.ctypes.data,
                    n, 9, None) == 0
    compare(y, ORC.mfi(hh, hl, hc, hv, 9), tol=0.0, min_exact=1.0, name="mfi host")
    x = _dev(hc[:40], dev)
    z = torch.zeros(40, device=dev)
    assert L.ct_ultosc_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), z.data_ptr(), 40, 7, 14, 40,
                         None) == 0
    torch.cuda.synchronize()
    assert torch.isnan(z).all()
    assert L.ct_ultosc_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), z.data_ptr(), 40, 0, 14, 28,
                         None) == 4
    assert L.ct_mfi_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), z.data_ptr(), 40, 1,
                      None) == 4
