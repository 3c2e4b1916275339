"""Streaming API (include/tacuda_stream.h) over numpy arrays, mirroring the
reference's binding (bindings/python/tacuda/__init__.py streaming classes):
Stream for one series on the host, StreamBatch for a portfolio whose
appends run on the host or on the GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from . import call, lib

# family -> (inputs read, number of outputs)
FAMILIES = {
    "sma": ("c", 1), "ema": ("c", 1), "wma": ("c", 1), "rsi": ("c", 1),
    "atr": ("hlc", 1), "natr": ("hlc", 1), "obv": ("cv", 1), "ad": ("hlcv", 1),
    "macd": ("c", 3), "bbands": ("c", 3),
}
EXEC = {"auto": 0, "cpu": 1, "gpu": 2}


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _host_inputs(family, inputs, arrs):
    """The family's host input arrays as contiguous float32, all of one
    length (the library reads every input for the length of the first)."""
    missing = [k for k in inputs if arrs[k] is None]
    if missing:
        raise ValueError(f"{family}: missing inputs {missing}")
    ins = [_f32(arrs[k]) for k in inputs]
    for a in ins[1:]:
        if a.shape != ins[0].shape:
            raise ValueError(f"{family}: input shapes differ: {ins[0].shape} vs {a.shape}")
    return ins


def _params(family, params):
    if family == "bbands":
        p, up, dn = (list(params) + [20, 2.0, 2.0][len(params):])[:3]
        return [int(p), float(up), float(dn)]
    return [int(v) for v in params]


class _Base:
    _destroy = ""

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            getattr(lib(), self._destroy)(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _q(self, name, ctype=ctypes.c_int):
        v = ctype()
        call(name, self._h, ctypes.byref(v))
        return v.value


class Stream(_Base):
    """Incremental state of one indicator over one series (host, FP64)."""

    _destroy = "ct_stream_destroy"

    def __init__(self, family, *params):
        self.family = family
        self.inputs, self.nout = FAMILIES[family]
        self._h = ctypes.c_void_p()
        call(f"ct_{family}_stream_create", *_params(family, params), ctypes.byref(self._h))

    def append(self, high=None, low=None, close=None, volume=None):
        arrs = {"h": high, "l": low, "c": close, "v": volume}
        ins = [a.ravel() for a in _host_inputs(self.family, self.inputs, arrs)]
        n = len(ins[0])
        outs = [np.empty(n, np.float32) for _ in range(self.nout)]
        args = [self._h] + [a.ctypes.data for a in ins] + [n] + [o.ctypes.data for o in outs]
        call(f"ct_{self.family}_stream_append", *args)
        return outs[0] if self.nout == 1 else tuple(outs)

    def reset(self):
        call("ct_stream_reset", self._h)

    @property
    def lookback(self):
        return self._q("ct_stream_lookback")

    @property
    def outputs(self):
        return self._q("ct_stream_outputs")

    @property
    def bars(self):
        return self._q("ct_stream_bars", ctypes.c_longlong)


class StreamBatch(_Base):
    """numSeries independent streams of one indicator; inputs and outputs
    are series-major [series, n] matrices."""

    _destroy = "ct_stream_batch_destroy"

    def __init__(self, family, num_series, *params, exec="auto"):
        self.family = family
        self.inputs, self.nout = FAMILIES[family]
        self.rows = int(num_series)
        self._h = ctypes.c_void_p()
        call(f"ct_{family}_stream_batch_create", *_params(family, params), self.rows,
             ctypes.byref(self._h))
        self.set_exec(exec)

    def set_exec(self, mode):
        call("ct_stream_batch_set_exec", self._h, EXEC[mode])

    def append(self, high=None, low=None, close=None, volume=None):
        arrs = {"h": high, "l": low, "c": close, "v": volume}
        ins = _host_inputs(self.family, self.inputs, arrs)
        if ins[0].size % self.rows:
            raise ValueError(f"{self.family}: {ins[0].size} values do not split into "
                             f"{self.rows} series")
        ins = [a.reshape(self.rows, -1) for a in ins]
        n = ins[0].shape[1]
        outs = [np.empty((self.rows, n), np.float32) for _ in range(self.nout)]
        args = [self._h] + [a.ctypes.data for a in ins] + [n] + [o.ctypes.data for o in outs]
        call(f"ct_{self.family}_stream_batch_append", *args)
        return outs[0] if self.nout == 1 else tuple(outs)

    def append_d(self, high=None, low=None, close=None, volume=None, outs=None):
        """Device tensors in and out (ctx_stream_batch_append_d) on the
        current stream; the state stays in HBM."""
        import torch

        arrs = {"h": high, "l": low, "c": close, "v": volume}
        ins = []
        for k in self.inputs:
            t = arrs[k]
            if t is None:
                raise ValueError(f"{self.family}: missing input {k!r}")
            ins.append(t)
        ref = ins[0]
        if ref.numel() == 0 or ref.numel() % self.rows:
            raise ValueError(f"{self.family}: {ref.numel()} values do not split into "
                             f"{self.rows} series")
        for t in ins[1:] + list(outs or []):
            if t.numel() != ref.numel() or t.device != ref.device:
                raise ValueError(f"{self.family}: inputs and outputs must hold {ref.numel()} "
                                 "values on one device")
        for t in ins + list(outs or []):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise ValueError("append_d expects contiguous float32 tensors on the GPU")
        n = ref.numel() // self.rows
        if outs is None:
            outs = [torch.empty_like(ref) for _ in range(self.nout)]
        elif len(outs) != self.nout:
            raise ValueError(f"{self.family}: expected {self.nout} output tensors")
        ptr = [t.data_ptr() if t is not None else None for t in (high, low, close, volume)]
        op = [o.data_ptr() for o in outs] + [None] * (3 - len(outs))
        call("ctx_stream_batch_append_d", self._h, *ptr, n, *op,
             torch.cuda.current_stream().cuda_stream)
        return outs[0] if self.nout == 1 else tuple(outs)

    def reset(self):
        call("ct_stream_batch_reset", self._h)

    @property
    def lookback(self):
        return self._q("ct_stream_batch_lookback")

    @property
    def outputs(self):
        return self._q("ct_stream_batch_outputs")

    @property
    def series(self):
        return self._q("ct_stream_batch_series")

    @property
    def bars(self):
        return self._q("ct_stream_batch_bars", ctypes.c_longlong)
