"""CPU oracle for the indicator core (test infrastructure only).

``oracle/oracle.c`` restates the TA-Lib 0.8.3 algorithms the reference vendors
(``third_party/ta_lib_0_8_3``) plus the reference's own semantics for the
families it does not vendor; ``oracle/candles.c`` restates the 61 TA-Lib
candlestick predicates.  ``oracle/_ref/libtaref.so`` is the reference's own
host build of TA-Lib 0.8.3 (``src/compat/Talib08Formulas.h``), used only to
check the restatement.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package.  Product code never does.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB: Optional[ctypes.CDLL] = None
_REF: Optional[ctypes.CDLL] = None

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.c_int
_D = ctypes.c_double


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        _LIB = ctypes.CDLL(path)
    return _LIB


def ref() -> Optional[ctypes.CDLL]:
    """The reference's TA-Lib host build, or None when it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libtaref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _call(fn, ins, nout, n, *params):
    ins = [_f32(a) for a in ins]
    outs = [np.empty(n, dtype=np.float32) for _ in range(nout)]
    args = [_fp(a) for a in ins] + [_fp(o) for o in outs] + [n] + list(params)
    rc = fn(*args)
    if rc != 0:
        raise ValueError(f"{fn.__name__} returned status {rc}")
    return outs[0] if nout == 1 else tuple(outs)


def _sig(fn, nin, nout, params):
    fn.argtypes = [_F] * (nin + nout) + [_I] + params
    fn.restype = _I
    return fn


class _Family:
    """Binds a C entry point family (``orc_*`` or ``ref_*``) to numpy."""

    def __init__(self, getlib, prefix):
        self._get = getlib
        self._prefix = prefix

    def _fn(self, name, nin, nout, params):
        L = self._get()
        if L is None:
            raise RuntimeError("library not available")
        return _sig(getattr(L, self._prefix + name), nin, nout, params)

    # single input, one period
    def sma(self, x, p):
        return _call(self._fn("sma", 1, 1, [_I]), [x], 1, len(x), p)

    def sum(self, x, p):
        return _call(self._fn("sum", 1, 1, [_I]), [x], 1, len(x), p)

    def wma(self, x, p):
        return _call(self._fn("wma", 1, 1, [_I]), [x], 1, len(x), p)

    def ema(self, x, p):
        return _call(self._fn("ema", 1, 1, [_I]), [x], 1, len(x), p)

    def dema(self, x, p):
        return _call(self._fn("dema", 1, 1, [_I]), [x], 1, len(x), p)

    def tema(self, x, p):
        return _call(self._fn("tema", 1, 1, [_I]), [x], 1, len(x), p)

    def t3(self, x, p, v):
        return _call(self._fn("t3", 1, 1, [_I, _D]), [x], 1, len(x), p, v)

    def kama(self, x, p, fast=2, slow=30):
        if self._prefix == "ref_":
            return _call(self._fn("kama", 1, 1, [_I]), [x], 1, len(x), p)
        return _call(self._fn("kama", 1, 1, [_I, _I, _I]), [x], 1, len(x), p, fast, slow)

    def rsi(self, x, p):
        return _call(self._fn("rsi", 1, 1, [_I]), [x], 1, len(x), p)

    def rma(self, x, p):
        return _call(self._fn("rma", 1, 1, [_I]), [x], 1, len(x), p)

    def var(self, x, p):
        return _call(self._fn("var", 1, 1, [_I]), [x], 1, len(x), p)

    def stddev(self, x, p):
        return _call(self._fn("stddev", 1, 1, [_I]), [x], 1, len(x), p)

    def min(self, x, p):
        return _call(self._fn("min", 1, 1, [_I]), [x], 1, len(x), p)

    def max(self, x, p):
        return _call(self._fn("max", 1, 1, [_I]), [x], 1, len(x), p)

    def macd(self, x, fast, slow, sig):
        return _call(self._fn("macd", 1, 3, [_I, _I, _I]), [x], 3, len(x), fast, slow, sig)

    def bbands(self, x, p, up, dn):
        return _call(self._fn("bbands", 1, 3, [_I, _D, _D]), [x], 3, len(x), p, up, dn)

    def atr(self, h, l, c, p):
        return _call(self._fn("atr", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def natr(self, h, l, c, p):
        return _call(self._fn("natr", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def trange(self, h, l, c):
        return _call(self._fn("trange", 3, 1, []), [h, l, c], 1, len(c))

    def stoch(self, h, l, c, kp, dp):
        if self._prefix == "ref_":
            return _call(self._fn("stoch", 3, 2, [_I, _I, _I]), [h, l, c], 2, len(c), kp, dp, dp)
        return _call(self._fn("stoch", 3, 2, [_I, _I]), [h, l, c], 2, len(c), kp, dp)

    def willr(self, h, l, c, p):
        return _call(self._fn("willr", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def cci(self, h, l, c, p):
        return _call(self._fn("cci", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def aroon(self, h, l, p):
        """Returns (down, up) like the reference ct_aroon_d."""
        return _call(self._fn("aroon", 2, 2, [_I]), [h, l], 2, len(h), p)

    def aroonosc(self, h, l, p):
        return _call(self._fn("aroonosc", 2, 1, [_I]), [h, l], 1, len(h), p)

    def plus_dm(self, h, l, p):
        return _call(self._fn("plus_dm", 2, 1, [_I]), [h, l], 1, len(h), p)

    def minus_dm(self, h, l, p):
        return _call(self._fn("minus_dm", 2, 1, [_I]), [h, l], 1, len(h), p)

    def plus_di(self, h, l, c, p):
        return _call(self._fn("plus_di", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def minus_di(self, h, l, c, p):
        return _call(self._fn("minus_di", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def dx(self, h, l, c, p):
        return _call(self._fn("dx", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def adx(self, h, l, c, p):
        return _call(self._fn("adx", 3, 1, [_I]), [h, l, c], 1, len(c), p)

    def obv(self, c, v):
        return _call(self._fn("obv", 2, 1, []), [c, v], 1, len(c))

    def ad(self, h, l, c, v):
        return _call(self._fn("ad", 4, 1, []), [h, l, c, v], 1, len(c))

    # price transforms, momentum, window midpoints, linear regression
    def avgprice(self, o, h, l, c):
        return _call(self._fn("avgprice", 4, 1, []), [o, h, l, c], 1, len(c))

    def medprice(self, h, l):
        return _call(self._fn("medprice", 2, 1, []), [h, l], 1, len(h))

    def typprice(self, h, l, c):
        return _call(self._fn("typprice", 3, 1, []), [h, l, c], 1, len(c))

    def wclprice(self, h, l, c):
        return _call(self._fn("wclprice", 3, 1, []), [h, l, c], 1, len(c))

    def momentum(self, x, p):
        return _call(self._fn("momentum", 1, 1, [_I]), [x], 1, len(x), p)

    def roc(self, x, p):
        return _call(self._fn("roc", 1, 1, [_I]), [x], 1, len(x), p)

    def rocp(self, x, p):
        return _call(self._fn("rocp", 1, 1, [_I]), [x], 1, len(x), p)

    def rocr(self, x, p):
        return _call(self._fn("rocr", 1, 1, [_I]), [x], 1, len(x), p)

    def rocr100(self, x, p):
        return _call(self._fn("rocr100", 1, 1, [_I]), [x], 1, len(x), p)

    def midprice(self, h, l, p):
        return _call(self._fn("midprice", 2, 1, [_I]), [h, l], 1, len(h), p)

    def midpoint(self, x, p):
        return _call(self._fn("midpoint", 1, 1, [_I]), [x], 1, len(x), p)

    def linearreg(self, x, p):
        return _call(self._fn("linearreg", 1, 1, [_I]), [x], 1, len(x), p)

    def linearreg_slope(self, x, p):
        return _call(self._fn("linearreg_slope", 1, 1, [_I]), [x], 1, len(x), p)

    def linearreg_intercept(self, x, p):
        return _call(self._fn("linearreg_intercept", 1, 1, [_I]), [x], 1, len(x), p)

    def linearreg_angle(self, x, p):
        return _call(self._fn("linearreg_angle", 1, 1, [_I]), [x], 1, len(x), p)

    def tsf(self, x, p):
        return _call(self._fn("tsf", 1, 1, [_I]), [x], 1, len(x), p)

    # element-wise math: op index into MATH1 / MATH2
    MATH1 = ("acos", "asin", "atan", "ceil", "cos", "cosh", "exp", "floor", "ln", "log10", "sin",
             "sinh", "sqrt", "tan", "tanh")
    MATH2 = ("add", "sub", "mult", "div")

    def math1(self, name, x):
        fn = getattr(self._get(), self._prefix + "math1")
        fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        fn.restype = ctypes.c_int
        x = _f32(x)
        out = np.empty(len(x), np.float32)
        rc = fn(self.MATH1.index(name), _fp(x), _fp(out), len(x))
        if rc != 0:
            raise ValueError(f"math1 {name} returned status {rc}")
        return out

    def math2(self, name, a, b):
        fn = getattr(self._get(), self._prefix + "math2")
        fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_int]
        fn.restype = ctypes.c_int
        a, b = _f32(a), _f32(b)
        out = np.empty(len(a), np.float32)
        rc = fn(self.MATH2.index(name), _fp(a), _fp(b), _fp(out), len(a))
        if rc != 0:
            raise ValueError(f"math2 {name} returned status {rc}")
        return out

    def bop(self, o, h, l, c):
        return _call(self._fn("bop", 4, 1, []), [o, h, l, c], 1, len(c))

    def ibs(self, h, l, c):
        return _call(self._fn("ibs", 3, 1, []), [h, l, c], 1, len(c))

    def change(self, x, p):
        return _call(self._fn("change", 1, 1, [_I]), [x], 1, len(x), p)

    def avgdev(self, x, p):
        return _call(self._fn("avgdev", 1, 1, [_I]), [x], 1, len(x), p)

    def minmax(self, x, p):
        return _call(self._fn("minmax", 1, 2, [_I]), [x], 2, len(x), p)

    # TA-Lib 0.8.3 row functions (orc_rows / ref_rows): name, inputs, outputs
    ROWS = {"cumsum": (1, 1), "pvt": (2, 1), "rma": (1, 1), "zlema": (1, 1), "efi": (2, 1),
            "vwma": (2, 1), "er": (1, 1), "vhf": (1, 1), "dpo": (1, 1), "chop": (3, 1),
            "adr": (2, 1), "ao": (2, 1), "qstick": (2, 1), "donchian": (2, 3),
            "marketfi": (3, 1), "wad": (3, 1), "mcgd": (1, 1), "rvol": (1, 1), "cmf": (4, 1),
            "emv": (3, 1), "fosc": (1, 1), "eri": (3, 2), "percentrank": (1, 1), "si": (4, 1),
            "asi_wilder": (4, 1), "vortex": (3, 2), "cmou": (1, 1)}

    def rows(self, name, ins, p0=0, p1=0, d0=0.0):
        """One row through the backend's leading-NaN rule; a tuple for
        several outputs."""
        nin, nout = self.ROWS[name]
        assert len(ins) == nin
        fn = getattr(self._get(), self._prefix + "rows")
        fn.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                       ctypes.c_int, ctypes.c_int, ctypes.c_double]
        fn.restype = ctypes.c_int
        ins = [_f32(a) for a in ins]
        n = len(ins[0])
        outs = [np.empty(n, np.float32) for _ in range(nout)]
        pin = (ctypes.c_void_p * 4)(*([a.ctypes.data for a in ins] + [None] * (4 - nin)))
        pout = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - nout)))
        rc = fn(name.encode(), ctypes.cast(pin, ctypes.c_void_p), ctypes.cast(pout, ctypes.c_void_p),
                n, int(p0), int(p1), float(d0))
        if rc != 0:
            raise ValueError(f"rows {name} returned status {rc}")
        return outs[0] if nout == 1 else tuple(outs)

    def ultosc(self, h, l, c, p1, p2, p3):
        return _call(self._fn("ultosc", 3, 1, [_I, _I, _I]), [h, l, c], 1, len(c), p1, p2, p3)

    def mfi(self, h, l, c, v, p):
        return _call(self._fn("mfi", 4, 1, [_I]), [h, l, c, v], 1, len(c), p)

    def skew(self, x, p):
        return _call(self._fn("skew", 1, 1, [_I]), [x], 1, len(x), p)

    def trendscore(self, x, lb):
        return _call(self._fn("trendscore", 1, 1, [_I]), [x], 1, len(x), lb)

    def market_strength(self, x, p):
        return _call(self._fn("market_strength", 1, 1, [_I]), [x], 1, len(x), p)

    def pivot(self, high, x, s):
        fn = getattr(self._get(), self._prefix + "pivot")
        fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        fn.restype = ctypes.c_int
        x = _f32(x)
        out = np.empty(len(x), np.float32)
        rc = fn(1 if high else 0, _fp(x), _fp(out), len(x), int(s))
        if rc != 0:
            raise ValueError(f"pivot returned status {rc}")
        return out

    ROLLING = ("mean", "max", "min", "std")

    def rolling(self, kind, x, p):
        fn = getattr(self._get(), self._prefix + "rolling")
        fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        fn.restype = ctypes.c_int
        x = _f32(x)
        out = np.empty(len(x), np.float32)
        rc = fn(self.ROLLING.index(kind), _fp(x), _fp(out), len(x), int(p))
        if rc != 0:
            raise ValueError(f"rolling {kind} returned status {rc}")
        return out

    def candle(self, name, o, h, l, c):
        """int32 signal of one of the 61 TA-Lib patterns (see candles.c)."""
        L = self._get()
        fn = L.orc_cdl
        fn.argtypes = [ctypes.c_char_p, _F, _F, _F, _F, ctypes.POINTER(ctypes.c_int32), _I]
        fn.restype = _I
        o, h, l, c = (_f32(a) for a in (o, h, l, c))
        out = np.empty(len(c), dtype=np.int32)
        rc = fn(name.encode(), _fp(o), _fp(h), _fp(l), _fp(c),
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(c))
        if rc != 0:
            raise ValueError(f"orc_cdl({name}) returned {rc}")
        return out


    def aroon_host(self, h, l, up_p, dn_p):
        """The host AROON call with two periods (reference api.cpp:2511-2570,
        Aroon.cu:12-66): (up, down, osc); up from orc_aroon over up_p, down
        over dn_p, osc = up - down (float32) from max(up_p, dn_p); every line
        NaN when either period exceeds the series.  pinned: each line equals
        the one-period oracle for its own period (tests/test_oracle.py)."""
        n = len(h)
        if up_p > n or dn_p > n:
            nan = np.full(n, np.nan, np.float32)
            return nan, nan.copy(), nan.copy()
        up = self.aroon(h, l, up_p)[1]
        dn = self.aroon(h, l, dn_p)[0]
        osc = np.full(n, np.nan, np.float32)
        m = max(up_p, dn_p)
        osc[m:] = up[m:] - dn[m:]
        return up, dn, osc

    def engulfing_legacy(self, o, h, l, c):
        """The reference's ct_cdl_bullish_engulfing / bearish_engulfing
        (CandleUtils.h:233-260 through BullishEngulfing.cu:10-23): TA-Lib's
        signed engulfing signal, from bar 1 on (the 61-pattern engulfing
        starts at bar 2).  From bar 2 it is the engulfing oracle itself; bar
        1 applies the same rule.  pinned: the engulfing oracle's own checks
        (tests/test_oracle_candles.py)."""
        o, c = _f32(o), _f32(c)
        out = self.candle("engulfing", o, h, l, c).astype(np.float32)
        out[:2] = 0.0
        if len(c) > 1:
            co, cc, po, pc = o[1], c[1], o[0], c[0]
            cur = 1 if cc >= co else -1
            prev = 1 if pc >= po else -1
            bull = cur == 1 and prev == -1 and ((cc >= po and co < pc) or (cc > po and co <= pc))
            bear = cur == -1 and prev == 1 and ((co >= pc and cc < po) or (co > pc and cc <= po))
            if bull or bear:
                out[1] = cur * (100 if (co != pc and cc != po) else 80)
        return out


    def trima(self, x, p):
        if self._prefix == "ref_":
            return self.ma(x, p, 5)
        return _call(self._fn("trima", 1, 1, [_I]), [x], 1, len(x), p)

    def macdext(self, x, fast, ft, slow, st, sig, sgt):
        """(macd, signal, hist) with a moving-average type per line
        (TA_MAType numbering)."""
        return _call(self._fn("macdext", 1, 3, [_I] * 6), [x], 3, len(x), fast, ft, slow, st,
                     sig, sgt)

    # -- neighbours only the vendored library has (prefix "ref_") ----------
    def donchian(self, h, l, p):
        """(upper, lower) bands: highest high and lowest low over p bars."""
        return _call(self._fn("donchian", 2, 2, [_I]), [h, l], 2, len(h), p)

    def stochf(self, h, l, c, kp, dp):
        """(fast %K, fast %D with SMA smoothing)."""
        return _call(self._fn("stochf", 3, 2, [_I, _I]), [h, l, c], 2, len(c), kp, dp)

    def ma(self, x, p, ma_type):
        return _call(self._fn("ma", 1, 1, [_I, _I]), [x], 1, len(x), p, ma_type)


ORC = _Family(lib, "orc_")
REF = _Family(ref, "ref_")
