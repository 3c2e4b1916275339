"""Price transforms, momentum family, window midpoints and the
linear-regression family on the GPU against the oracle: device single calls
(ct_<name>_d and the _device aliases), dense batches (ct_<name>_batch_d on
ragged-length-free matrices of odd shapes), host single and batch calls,
non-finite and zero inputs, period edge cases and argument errors."""
import ctypes

import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ONE = ["momentum", "roc", "rocp", "rocr", "rocr100", "midpoint", "linearreg", "linearreg_slope",
       "linearreg_intercept", "linearreg_angle", "tsf"]
DEVICE_ALIAS = ["momentum", "roc", "rocp", "rocr", "rocr100"]
PRICE = {"avgprice": ("o", "h", "l", "c"), "medprice": ("h", "l"), "typprice": ("h", "l", "c"),
         "wclprice": ("h", "l", "c")}


def _tol(name):
    # atan on the device and in libm may round the last double bit apart
    return (1e-6, 0.999) if name == "linearreg_angle" else (0.0, 1.0)


def _check(got, want, name):
    tol, ex = _tol(name)
    compare(got, want, tol=tol, min_exact=ex, name=name)


def _oracle(name, d, p=None):
    from oracle import ORC

    if name in PRICE:
        return getattr(ORC, name)(*[d[{"o": "open", "h": "high", "l": "low", "c": "close"}[k]]
                                    for k in PRICE[name]])
    if name == "midprice":
        return ORC.midprice(d["high"], d["low"], p)
    return getattr(ORC, name)(d["close"], p)


def _data(n, rows=1, seed=0):
    """Seeded OHLC; with rows > 1 every array is [rows, n]."""
    import tacuda_amd as ta

    d = ta.synth_ohlcv(n, rows=rows, seed=seed)
    return d if rows == 1 else {k: np.asarray(v).reshape(rows, n) for k, v in d.items()}


def _row(d, r, rows):
    return {k: (v if rows == 1 else v[r]) for k, v in d.items()}


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("name", ONE)
@pytest.mark.parametrize("p", [1, 5, 14])
def test_single_device_calls(dev, name, p):
    import tacuda_amd as ta

    print(f"transforms device {name} p{p}", flush=True)
    L = ta.lib()
    d = _data(3001, seed=21)
    x = d["close"].copy()
    x[1500] = 0.0  # a zero earlier price for the ratios
    x[2000] = np.nan  # a hole
    d = dict(d, close=x)
    xi = _dev(x, dev)
    y = torch.empty_like(xi)
    rc = getattr(L, f"ct_{name}_d")(xi.data_ptr(), y.data_ptr(), len(x), p, None)
    assert rc == 0, L.ct_get_last_error()
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), _oracle(name, d, p), name)
    if name in DEVICE_ALIAS:
        y2 = torch.empty_like(xi)
        assert getattr(L, f"ct_{name}_device")(xi.data_ptr(), y2.data_ptr(), len(x), p, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(y2), torch.nan_to_num(y))


@pytest.mark.parametrize("name", ONE + ["midprice"])
@pytest.mark.parametrize("shape", [(7, 777), (40, 234), (1, 65)])
def test_dense_batches(dev, name, shape):
    import tacuda_amd as ta

    rows, n = shape
    p = 9
    print(f"transforms batch {name} {rows}x{n}", flush=True)
    L = ta.lib()
    d = _data(n, rows=rows, seed=23)
    y = torch.empty((rows, n), dtype=torch.float32, device=dev)
    if name == "midprice":
        h, l = _dev(d["high"], dev), _dev(d["low"], dev)
        rc = L.ct_midprice_batch_d(h.data_ptr(), l.data_ptr(), y.data_ptr(), n, rows, p, 4, None)
    else:
        x = _dev(d["close"], dev)
        rc = getattr(L, f"ct_{name}_batch_d")(x.data_ptr(), y.data_ptr(), n, rows, p, 4, None)
    assert rc == 0, L.ct_get_last_error()
    got = y.cpu().numpy()
    for r in range(rows):
        _check(got[r], _oracle(name, _row(d, r, rows), p), f"{name} row {r}")


@pytest.mark.parametrize("name", list(PRICE))
def test_price_transforms_device_and_batch(dev, name):
    import tacuda_amd as ta

    print(f"transforms price {name}", flush=True)
    L = ta.lib()
    rows, n = 6, 1003
    d = _data(n, rows=rows, seed=25)
    keys = {"o": "open", "h": "high", "l": "low", "c": "close"}
    ins = [_dev(d[keys[k]], dev) for k in PRICE[name]]
    y = torch.empty((rows, n), dtype=torch.float32, device=dev)
    rc = getattr(L, f"ct_{name}_batch_d")(*[t.data_ptr() for t in ins], y.data_ptr(), n, rows, 1,
                                          None)
    assert rc == 0, L.ct_get_last_error()
    got = y.cpu().numpy()
    for r in range(rows):
        _check(got[r], _oracle(name, {k: v[r] for k, v in d.items()}), f"{name} row {r}")
    y1 = torch.empty(n, dtype=torch.float32, device=dev)
    rc = getattr(L, f"ct_{name}_d")(*[t[0].data_ptr() for t in ins], y1.data_ptr(), n, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(y1.cpu().numpy(), got[0])


def _host_call(L, fn, arrays, out, *params):
    args = [a.ctypes.data for a in arrays] + [out.ctypes.data] + list(params) + [None]
    rc = getattr(L, fn)(*args)
    assert rc == 0, (fn, L.ct_get_last_error())


@pytest.mark.parametrize("name", ONE + ["midprice"] + list(PRICE))
def test_host_single_and_batch(dev, name):
    import tacuda_amd as ta

    print(f"transforms host {name}", flush=True)
    L = ta.lib()
    rows, n, p = 3, 500, 6
    d = _data(n, rows=rows, seed=27)
    keys = {"o": "open", "h": "high", "l": "low", "c": "close"}
    if name in PRICE:
        ins, params = [np.ascontiguousarray(d[keys[k]]) for k in PRICE[name]], ()
    elif name == "midprice":
        ins, params = [np.ascontiguousarray(d["high"]), np.ascontiguousarray(d["low"])], (p,)
    else:
        ins, params = [np.ascontiguousarray(d["close"])], (p,)
    out = np.empty((rows, n), np.float32)
    _host_call(L, f"ct_{name}_batch", ins, out, n, rows, *params, 4)
    one = np.empty(n, np.float32)
    _host_call(L, f"ct_{name}", [np.ascontiguousarray(a[1]) for a in ins], one, n, *params)
    for r in range(rows):
        _check(out[r], _oracle(name, {k: v[r] for k, v in d.items()}, p), f"{name} host row {r}")
    assert np.array_equal(np.nan_to_num(one), np.nan_to_num(out[1]))


def test_period_edges_and_errors(dev):
    import tacuda_amd as ta

    print("transforms edges", flush=True)
    L = ta.lib()
    x = _dev(_data(50, seed=29)["close"], dev)
    y = torch.zeros_like(x)
    # a period reaching past the series: success, all NaN (MOMENTUM/ROC*: at
    # the series length already)
    for name, p in (("momentum", 50), ("roc", 60), ("midpoint", 51), ("linearreg", 51),
                    ("tsf", 80)):
        assert getattr(L, f"ct_{name}_d")(x.data_ptr(), y.data_ptr(), 50, p, None) == 0, name
        torch.cuda.synchronize()
        assert torch.isnan(y).all(), name
        y.zero_()
    # period 0 and a null pointer are invalid arguments
    assert L.ct_roc_d(x.data_ptr(), y.data_ptr(), 50, 0, None) == 4
    assert L.ct_linearreg_batch_d(None, y.data_ptr(), 50, 1, 5, 1, None) == 4
    assert b"null" in L.ct_get_last_error()
    assert L.ct_medprice_d(x.data_ptr(), None, y.data_ptr(), 50, None) == 4
    assert L.ct_typprice_batch_d(x.data_ptr(), x.data_ptr(), x.data_ptr(), y.data_ptr(), 0, 1, 1,
                                 None) == 4


def test_midpoint_leading_nan_and_midprice_nan(dev):
    """A MIDPOINT window that starts in the leading NaN run is NaN; interior
    NaN samples are skipped by both midpoints."""
    import tacuda_amd as ta
    from oracle import ORC

    print("transforms midpoint NaN", flush=True)
    L = ta.lib()
    d = _data(400, seed=31)
    x = d["close"].copy()
    x[:9] = np.nan
    x[200] = np.nan
    h = d["high"].copy()
    h[150] = np.nan
    xi, hi, li = _dev(x, dev), _dev(h, dev), _dev(d["low"], dev)
    y = torch.empty_like(xi)
    assert L.ct_midpoint_d(xi.data_ptr(), y.data_ptr(), 400, 7, None) == 0
    torch.cuda.synchronize()
    compare(y.cpu().numpy(), ORC.midpoint(x, 7), tol=0.0, min_exact=1.0, name="midpoint nan")
    assert L.ct_midprice_d(hi.data_ptr(), li.data_ptr(), y.data_ptr(), 400, 7, None) == 0
    torch.cuda.synchronize()
    compare(y.cpu().numpy(), ORC.midprice(h, d["low"], 7), tol=0.0, min_exact=1.0,
            name="midprice nan")
