"""TA-Lib 0.8.3 row functions (CUMSUM, PVT, RMA, ZLEMA, EFI, VWMA, ER, VHF,
DPO, CHOP, ADR, AO, QSTICK, DONCHIAN, MARKETFI, WAD, MCGD, RVOL, CMF, EMV,
FOSC, ERI, PERCENTRANK, SI, ASI_WILDER, VORTEX, CMOU) on the GPU
against the oracle, bit for bit: device single calls on long rows (one wave
per row), dense batches of many short rows (one lane per row) with leading
and interior NaNs, zero-volume and unchanged-price runs, host single and
batch calls, period edges and argument errors.  CHOP takes log10 in double,
where the device library and the host's libm may round the last double bit
apart: at least 99.9% of its outputs must be bit-identical and the rest
within 1e-6 relative."""
import ctypes

import numpy as np
import pytest

from tests.parity import compare
from tests.test_oracle_rows import NPAR, SPEC, bars, inputs, params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NOUT = {"donchian": 3, "eri": 2, "vortex": 2}


def _check(name, got, want, what):
    tol, ex = (1e-6, 0.999) if name == "chop" else (0.0, 1.0)
    compare(got, want, tol=tol, min_exact=ex, name=what)


def _want(name, ins, p):
    from oracle import ORC

    w = ORC.rows(name, ins, *p)
    return w if isinstance(w, tuple) else (w,)


def _call_args(name, p):
    k = NPAR.get(name, 1)
    if k == "d":
        return [ctypes.c_double(p[2])]
    if k == "1d":
        return [p[0], ctypes.c_double(p[2])]
    return list(p[:k])


@pytest.mark.parametrize("name", list(SPEC))
def test_single_long_row(dev, name):
    import tacuda_amd as ta

    L = ta.lib()
    d = bars(20000, 5, holes=True)
    ins = inputs(name, d)
    nout = NOUT.get(name, 1)
    for p in params(name)[1:3] or params(name):
        print(f"rows {name} single p{p}", flush=True)
        xi = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in ins]
        ys = [torch.zeros(20000, device=dev) for _ in range(nout)]
        rc = getattr(L, f"ct_{name}_d")(*[t.data_ptr() for t in xi], *[y.data_ptr() for y in ys],
                                        20000, *_call_args(name, p), None)
        assert rc == 0, L.ct_get_last_error()
        torch.cuda.synchronize()
        for k, w in enumerate(_want(name, ins, p)):
            _check(name, ys[k].cpu().numpy(), w, f"{name}{p} out {k}")


@pytest.mark.parametrize("name", list(SPEC))
@pytest.mark.parametrize("rows,n", [(300, 257), (7, 5000)])
def test_dense_batches(dev, name, rows, n):
    import tacuda_amd as ta

    L = ta.lib()
    p = params(name)[-1]
    print(f"rows {name} batch {rows}x{n} p{p}", flush=True)
    per_row = [bars(n, 100 + r, holes=(r % 3 == 0) and n > 300) for r in range(rows)]
    ins = [np.stack(col) for col in zip(*[inputs(name, d) for d in per_row])]
    nout = NOUT.get(name, 1)
    xi = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in ins]
    ys = [torch.zeros((rows, n), device=dev) for _ in range(nout)]
    rc = getattr(L, f"ct_{name}_batch_d")(*[t.data_ptr() for t in xi], *[y.data_ptr() for y in ys],
                                          n, rows, *_call_args(name, p), 4, None)
    assert rc == 0, L.ct_get_last_error()
    torch.cuda.synchronize()
    got = [y.cpu().numpy() for y in ys]
    for r in range(0, rows, max(1, rows // 40)):
        for k, w in enumerate(_want(name, [a[r] for a in ins], p)):
            _check(name, got[k][r], w, f"{name}{p} row {r} out {k}")
    # host batch form equals the device batch
    hb = [np.empty((rows, n), np.float32) for _ in range(nout)]
    rc = getattr(L, f"ct_{name}_batch")(*[np.ascontiguousarray(a).ctypes.data for a in ins],
                                        *[o.ctypes.data for o in hb], n, rows,
                                        *_call_args(name, p), 4, None)
    assert rc == 0, L.ct_get_last_error()
    for k in range(nout):
        assert np.array_equal(np.nan_to_num(hb[k], nan=-7.0), np.nan_to_num(got[k], nan=-7.0))


@pytest.mark.parametrize("name", list(SPEC))
def test_host_single(dev, name):
    import tacuda_amd as ta

    print(f"rows {name} host", flush=True)
    L = ta.lib()
    d = bars(3001, 8)
    ins = [np.ascontiguousarray(a) for a in inputs(name, d)]
    p = params(name)[-1]
    nout = NOUT.get(name, 1)
    outs = [np.empty(3001, np.float32) for _ in range(nout)]
    rc = getattr(L, f"ct_{name}")(*[a.ctypes.data for a in ins], *[o.ctypes.data for o in outs],
                                  3001, *_call_args(name, p), None)
    assert rc == 0, L.ct_get_last_error()
    for k, w in enumerate(_want(name, ins, p)):
        _check(name, outs[k], w, f"{name} host out {k}")


def test_row_edges_and_errors(dev):
    import tacuda_amd as ta

    print("rows edges", flush=True)
    L = ta.lib()
    x = torch.arange(1, 41, dtype=torch.float32, device=dev)
    y = torch.zeros_like(x)
    # a period beyond the series: success, all NaN
    for name, p in (("rma", 41), ("er", 40), ("vhf", 45), ("dpo", 100), ("mcgd", 41)):
        assert getattr(L, f"ct_{name}_d")(x.data_ptr(), y.data_ptr(), 40, p, None) == 0, name
        torch.cuda.synchronize()
        assert torch.isnan(y).all(), name
    # periods outside TA-Lib's ranges, null pointers, empty shapes
    assert L.ct_er_d(x.data_ptr(), y.data_ptr(), 40, 1, None) == 4
    assert L.ct_rma_d(x.data_ptr(), y.data_ptr(), 40, 100001, None) == 4
    assert L.ct_ao_d(x.data_ptr(), x.data_ptr(), y.data_ptr(), 40, 1, 5, None) == 4
    assert L.ct_donchian_d(x.data_ptr(), x.data_ptr(), y.data_ptr(), None, y.data_ptr(), 40, 5,
                           None) == 4
    assert L.ct_cumsum_batch_d(x.data_ptr(), y.data_ptr(), 40, 0, 1, None) == 4
    assert L.ct_wad(None, None, None, None, 40, None) == 4
    assert L.ct_percentrank_d(x.data_ptr(), y.data_ptr(), 40, 10001, None) == 4
    assert L.ct_emv_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), y.data_ptr(), 40, 5,
                      ctypes.c_double(0.5), None) == 4
    assert L.ct_si_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), y.data_ptr(), 40,
                     ctypes.c_double(0.0), None) == 4
