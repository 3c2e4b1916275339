"""Torch-tensor front end of the C ABI.

Each function takes CUDA (HIP) float32 tensors, allocates outputs with
``torch.empty_like`` and calls the matching ``ct_*_d`` (1-D input) or
``ct_*_batch_d`` (2-D series-major input) entry point on the current stream.
Names and argument meaning follow the reference's ctypes binding.
"""
from __future__ import annotations

import torch

import ctypes

from . import PACK_OUTPUTS, SCAN_OUTPUTS, PackOutputs, PackParams, ScanOutputs, ScanParams, call

P = ctypes.c_void_p

_ONE_IN = ("sma", "ema", "wma", "dema", "tema", "rsi", "stddev", "var", "sum", "max", "min")
_HLC = ("natr", "adx", "dx", "plus_di", "minus_di", "willr", "cci")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check(*ts):
    """Inputs of one call: contiguous float32 GPU tensors of one shape (the
    kernels index every input with the shape of the last one, so a shorter
    high or low would be read past its end; the reference binding rejects
    mismatched lengths too, bindings/python/tacuda/__init__.py:852,914)."""
    for t in ts:
        if t.dim() not in (1, 2):
            raise ValueError("expected a 1-D series or a 2-D [series, bars] matrix")
    for t in ts[1:]:
        if t.shape != ts[0].shape:
            raise ValueError(f"input shapes differ: {tuple(ts[0].shape)} vs {tuple(t.shape)}")
    for t in ts:
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise ValueError("expected contiguous float32 tensors on the GPU")
        if t.device != ts[0].device:
            raise ValueError("inputs are on different devices")


def _check_ragged(layout, *ts):
    """Ragged inputs: 1-D tensors holding exactly offsets[-1] bars."""
    total = int(layout.offsets[-1])
    for t in ts:
        if t.dim() != 1 or t.numel() != total:
            raise ValueError(f"ragged inputs must be 1-D with offsets[-1] = {total} bars, "
                             f"got shape {tuple(t.shape)}")
    _check(*ts)


def _shape(x):
    return (x.shape[0], 1) if x.dim() == 1 else (x.shape[1], x.shape[0])


def _dispatch(name, x, ins, outs, *params, batch_params=None):
    n, rows = _shape(x)
    if x.dim() == 1:
        call(f"ct_{name}_d", *[t.data_ptr() for t in ins + outs], n, *params, _stream())
    else:
        bp = params if batch_params is None else batch_params
        call(f"ct_{name}_batch_d", *[t.data_ptr() for t in ins + outs], n, rows, *bp, 0,
             _stream())


def _make_one(name):
    def f(x: torch.Tensor, period: int) -> torch.Tensor:
        _check(x)
        y = torch.empty_like(x)
        _dispatch(name, x, [x], [y], int(period))
        return y

    f.__name__ = name
    f.__doc__ = f"{name.upper()} of a series (ct_{name}_d / ct_{name}_batch_d)."
    return f


def _make_hlc(name):
    def f(high, low, close, period: int) -> torch.Tensor:
        _check(high, low, close)
        y = torch.empty_like(close)
        _dispatch(name, close, [high, low, close], [y], int(period))
        return y

    f.__name__ = name
    f.__doc__ = f"{name.upper()} of high/low/close (ct_{name}_d / ct_{name}_batch_d)."
    return f


for _n in _ONE_IN:
    globals()[_n] = _make_one(_n)
for _n in _HLC:
    globals()[_n] = _make_hlc(_n)


def t3(x, period: int, vfactor: float = 0.7):
    _check(x)
    y = torch.empty_like(x)
    _dispatch("t3", x, [x], [y], int(period), float(vfactor))
    return y


def kama(x, period: int, fast: int = 2, slow: int = 30):
    _check(x)
    y = torch.empty_like(x)
    _dispatch("kama", x, [x], [y], int(period), int(fast), int(slow))
    return y


def macd(x, fast: int = 12, slow: int = 26, signal: int = 9, ma_type: int = 1):
    _check(x)
    m, s, h = (torch.empty_like(x) for _ in range(3))
    _dispatch("macd", x, [x], [m, s, h], fast, slow, signal, ma_type)
    return m, s, h


def trima(x, period: int):
    """TA-Lib TRIMA (ct_trima_d / ct_trima_batch_d)."""
    _check(x)
    y = torch.empty_like(x)
    _dispatch("trima", x, [x], [y], int(period))
    return y


MA_TYPES = {"sma": 0, "ema": 1, "wma": 2, "dema": 3, "tema": 4, "trima": 5, "kama": 6, "t3": 8}


def macdext(x, fast: int = 12, fast_type="ema", slow: int = 26, slow_type="ema", signal: int = 9,
            signal_type="ema"):
    """TA-Lib MACDEXT: a moving-average type per line (names of MA_TYPES or
    ctMaType_t values), ct_macdext_d / ct_macdext_batch_d."""
    _check(x)
    t = [MA_TYPES.get(v, v) if isinstance(v, str) else int(v)
         for v in (fast_type, slow_type, signal_type)]
    m, s, h = (torch.empty_like(x) for _ in range(3))
    _dispatch("macdext", x, [x], [m, s, h], int(fast), t[0], int(slow), t[1], int(signal), t[2])
    return m, s, h


def bbands(x, period: int = 20, up: float = 2.0, dn: float = 2.0):
    _check(x)
    u, m, l = (torch.empty_like(x) for _ in range(3))
    _dispatch("bbands", x, [x], [u, m, l], int(period), float(up), float(dn))
    return u, m, l


def atr(high, low, close, period: int = 14):
    _check(high, low, close)
    y = torch.empty_like(close)
    _dispatch("atr", close, [high, low, close], [y], int(period), 0.0)
    return y


def trange(high, low, close):
    _check(high, low, close)
    y = torch.empty_like(close)
    _dispatch("trange", close, [high, low, close], [y])
    return y


def stochastic(high, low, close, k_period: int = 14, d_period: int = 3):
    _check(high, low, close)
    k, d = torch.empty_like(close), torch.empty_like(close)
    _dispatch("stochastic", close, [high, low, close], [k, d], int(k_period), int(d_period))
    return k, d


def aroon(high, low, period: int = 14):
    """Returns (down, up) like the reference ct_aroon_d."""
    _check(high, low)
    d, u = torch.empty_like(high), torch.empty_like(high)
    _dispatch("aroon", high, [high, low], [d, u], int(period))
    return d, u


def aroonosc(high, low, period: int = 14):
    _check(high, low)
    y = torch.empty_like(high)
    _dispatch("aroonosc", high, [high, low], [y], int(period))
    return y


def plus_dm(high, low, period: int = 14):
    _check(high, low)
    y = torch.empty_like(high)
    _dispatch("plus_dm", high, [high, low], [y], int(period))
    return y


def minus_dm(high, low, period: int = 14):
    _check(high, low)
    y = torch.empty_like(high)
    _dispatch("minus_dm", high, [high, low], [y], int(period))
    return y


PACK_DEFAULTS = dict(sma_period=20, ema_period=20, wma_period=20, dema_period=20, tema_period=20,
                     rsi_period=14, macd_fast=12, macd_slow=26, macd_signal=9, bb_period=20,
                     bb_up=2.0, bb_dn=2.0, atr_period=14, adx_period=14, stoch_k=14, stoch_d=3,
                     willr_period=14, cci_period=14)


def pack(high, low, close, outputs=PACK_OUTPUTS, **periods):
    """Several indicators of the same bars in one fused call (ctx_pack_batch_d).

    ``outputs`` names the ctxPackOutputs fields to compute; ``periods``
    overrides fields of ctxPackParams (defaults: PACK_DEFAULTS).  EMA, DEMA
    and TEMA share one period.  Returns a dict of tensors shaped like close.
    """
    if high is None:
        high = low = close
    _check(high, low, close)
    unknown = set(outputs) - set(PACK_OUTPUTS)
    if unknown:
        raise ValueError(f"unknown pack outputs {sorted(unknown)}")
    prm = PackParams(**{**PACK_DEFAULTS, **periods})
    res = {k: torch.empty_like(close) for k in outputs}
    outs = PackOutputs(**{k: v.data_ptr() for k, v in res.items()})
    n, rows = _shape(close)
    call("ctx_pack_batch_d", high.data_ptr(), low.data_ptr(), close.data_ptr(), n, rows, prm, outs,
         _stream())
    return res


SCAN_DEFAULTS = dict(ema_period=200, kama_period=30, kama_fast=2, kama_slow=30, t3_period=5,
                     t3_vfactor=0.7, sma_period=200)


def scan(close, outputs=SCAN_OUTPUTS, **periods):
    """EMA, KAMA, T3 and SMA of the same close rows from one fused launch
    sequence (ctx_scan_batch_d); ``outputs`` names the ones to compute,
    ``periods`` overrides SCAN_DEFAULTS.  Returns a dict of tensors shaped
    like close."""
    _check(close)
    unknown = set(outputs) - set(SCAN_OUTPUTS)
    if unknown:
        raise ValueError(f"unknown scan outputs {sorted(unknown)}")
    prm = ScanParams(**{**SCAN_DEFAULTS, **periods})
    res = {k: torch.empty_like(close) for k in outputs}
    outs = ScanOutputs(**{k: v.data_ptr() for k, v in res.items()})
    n, rows = _shape(close)
    call("ctx_scan_batch_d", close.data_ptr(), n, rows, prm, outs, _stream())
    return res


class Ragged:
    """Ragged batch layout (ctx_ragged_create): series r occupies
    [offsets[r], offsets[r+1]) of 1-D concatenated tensors."""

    def __init__(self, offsets):
        import ctypes

        import numpy as np

        from . import lib

        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.offsets = off
        self.rows = len(off) - 1
        h = ctypes.c_void_p()
        call("ctx_ragged_create", off.ctypes.data, self.rows, ctypes.byref(h))
        self._h = h
        self._lib = lib()

    def close(self):
        if self._h is not None and self._h.value:
            self._lib.ctx_ragged_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def one(self, name, x, *params):
        """Single-input family (sma, ema, ..., t3, kama) over the layout."""
        _check_ragged(self, x)
        y = torch.empty_like(x)
        call(f"ctx_{name}_ragged_d", x.data_ptr(), y.data_ptr(), self._h, *params, _stream())
        return y

    def hlc(self, name, high, low, close, *params):
        _check_ragged(self, high, low, close)
        y = torch.empty_like(close)
        call(f"ctx_{name}_ragged_d", high.data_ptr(), low.data_ptr(), close.data_ptr(),
             y.data_ptr(), self._h, *params, _stream())
        return y

    def macd(self, x, fast=12, slow=26, signal=9):
        _check_ragged(self, x)
        m, s, h = (torch.empty_like(x) for _ in range(3))
        call("ctx_macd_ragged_d", x.data_ptr(), m.data_ptr(), s.data_ptr(), h.data_ptr(), self._h,
             fast, slow, signal, _stream())
        return m, s, h

    def bbands(self, x, period=20, up=2.0, dn=2.0):
        _check_ragged(self, x)
        u, m, l = (torch.empty_like(x) for _ in range(3))
        call("ctx_bbands_ragged_d", x.data_ptr(), u.data_ptr(), m.data_ptr(), l.data_ptr(),
             self._h, period, up, dn, _stream())
        return u, m, l

    def stochastic(self, high, low, close, k_period=14, d_period=3):
        _check_ragged(self, high, low, close)
        k, d = torch.empty_like(close), torch.empty_like(close)
        call("ctx_stochastic_ragged_d", high.data_ptr(), low.data_ptr(), close.data_ptr(),
             k.data_ptr(), d.data_ptr(), self._h, k_period, d_period, _stream())
        return k, d

    def trange(self, high, low, close):
        _check_ragged(self, high, low, close)
        y = torch.empty_like(close)
        call("ctx_trange_ragged_d", high.data_ptr(), low.data_ptr(), close.data_ptr(),
             y.data_ptr(), self._h, _stream())
        return y

    def pack(self, high, low, close, outputs=PACK_OUTPUTS, **periods):
        """The fused multi-indicator launch over the layout (ctx_pack_ragged_d)."""
        if high is None:
            high = low = close
        _check_ragged(self, high, low, close)
        prm = PackParams(**{**PACK_DEFAULTS, **periods})
        res = {k: torch.empty_like(close) for k in outputs}
        outs = PackOutputs(**{k: v.data_ptr() for k, v in res.items()})
        call("ctx_pack_ragged_d", high.data_ptr(), low.data_ptr(), close.data_ptr(), self._h, prm,
             outs, _stream())
        return res


class Plan:
    """A tuned, graph-captured fused pass over fixed buffers
    (ctx_pack_plan_create).  Output tensors are allocated once and rewritten
    by every launch(); inputs are read from the given tensors at launch time."""

    def __init__(self, high, low, close, outputs=PACK_OUTPUTS, layout=None, tune=True,
                 **periods):
        import ctypes

        from . import lib

        if high is None:
            high = low = close
        if layout is None:
            _check(high, low, close)
        else:
            _check_ragged(layout, high, low, close)
        self.inputs = (high, low, close)
        self.outputs = {k: torch.empty_like(close) for k in outputs}
        prm = PackParams(**{**PACK_DEFAULTS, **periods})
        outs = PackOutputs(**{k: v.data_ptr() for k, v in self.outputs.items()})
        n, rows = _shape(close) if layout is None else (0, 0)
        h = ctypes.c_void_p()
        torch.cuda.current_stream().synchronize()
        call("ctx_pack_plan_create", high.data_ptr(), low.data_ptr(), close.data_ptr(), n, rows,
             layout.handle if layout is not None else None, prm, outs, int(tune),
             ctypes.byref(h))
        self._h = h
        self._lib = lib()
        self.layout = layout

    def launch(self):
        call("ctx_plan_launch", self._h, _stream())
        return self.outputs

    def info(self):
        import ctypes

        k, w = ctypes.c_int(), ctypes.c_int()
        tu, du = ctypes.c_float(), ctypes.c_float()
        call("ctx_plan_info", self._h, ctypes.byref(k), ctypes.byref(w), ctypes.byref(tu),
             ctypes.byref(du))
        return {"bars_per_lane": k.value, "waves_per_group": w.value, "tuned_us": tu.value,
                "default_us": du.value}

    def close(self):
        if self._h is not None and self._h.value:
            self._lib.ctx_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cdl_names():
    """The 61 TA-Lib pattern names, in the library's pattern order."""
    from . import lib

    L = lib()
    return [L.ctx_cdl_name(k).decode() for k in range(L.ctx_cdl_count())]


def cdl(name, open_, high, low, close):
    """int32 signal of one pattern (ctx_cdl_batch_d); 1-D or [series, bars]."""
    _check(open_, high, low, close)
    k = cdl_names().index(name)
    n, rows = _shape(close)
    out = torch.empty(close.shape, dtype=torch.int32, device=close.device)
    call("ctx_cdl_batch_d", k, open_.data_ptr(), high.data_ptr(), low.data_ptr(),
         close.data_ptr(), out.data_ptr(), n, rows, _stream())
    return out


def cdl_all(open_, high, low, close):
    """All 61 patterns in one launch: int32 tensor [61, *close.shape]."""
    _check(open_, high, low, close)
    n, rows = _shape(close)
    out = torch.empty((61,) + tuple(close.shape), dtype=torch.int32, device=close.device)
    call("ctx_cdl_all_batch_d", open_.data_ptr(), high.data_ptr(), low.data_ptr(),
         close.data_ptr(), out.data_ptr(), n, rows, _stream())
    return out


def cdl_legacy(name, open_, high, low, close):
    """The reference's float-signal call ct_cdl_<name>_d / _batch_d."""
    _check(open_, high, low, close)
    from . import lib

    L = lib()
    n, rows = _shape(close)
    out = torch.empty_like(close)
    args = [open_.data_ptr(), high.data_ptr(), low.data_ptr(), close.data_ptr(), out.data_ptr()]
    if close.dim() == 1:
        fn = getattr(L, f"ct_cdl_{name}_d")
        fn.argtypes = [P] * 5 + [ctypes.c_int, ctypes.c_void_p]
        rc = fn(*args, n, _stream())
    else:
        fn = getattr(L, f"ct_cdl_{name}_batch_d")
        fn.argtypes = [P] * 5 + [ctypes.c_int] * 3 + [ctypes.c_void_p]
        rc = fn(*args, n, rows, 0, _stream())
    if rc != 0:
        from . import TacudaError, last_error

        raise TacudaError(f"ct_cdl_{name}", rc, last_error())
    return out
