"""tacuda_amd: MI355X-native TACUDA indicator core.

The compute path is ``libtacuda_amd.so`` (hand-written HIP kernels for gfx950
behind the reference's C ABI).  This module binds it with ctypes, the way the
reference's Python binding does (``bindings/python/tacuda/__init__.py:46-123``),
and adds torch-tensor helpers that pass ``data_ptr()`` and the current HIP
stream.  There is no CPU fallback: every helper raises if the library is
missing.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TACUDA_AMD_LIB", os.path.join(_HERE, "libtacuda_amd.so"))

_lib: Optional[ctypes.CDLL] = None

c_int = ctypes.c_int
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
P = ctypes.c_void_p  # device or host float pointer

# name -> argument types (after which `stream` follows unless noted)
_SIGS = {}


def _sig(name, *args, stream=True, restype=c_int):
    _SIGS[name] = (list(args) + ([c_void_p] if stream else []), restype)


# device single series (include/tacuda_device.h)
for _n in ("sma", "ema", "wma", "dema", "tema", "rsi", "stddev", "var", "sum", "max", "min"):
    _sig(f"ct_{_n}_d", P, P, c_int, c_int)
_sig("ct_t3_d", P, P, c_int, c_int, c_float)
_sig("ct_kama_d", P, P, c_int, c_int, c_int, c_int)
_sig("ct_macd_d", P, P, P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ct_bbands_d", P, P, P, P, c_int, c_int, c_float, c_float)
_sig("ct_atr_d", P, P, P, P, c_int, c_int, c_float)
for _n in ("natr", "adx", "dx", "plus_di", "minus_di", "willr", "cci"):
    _sig(f"ct_{_n}_d", P, P, P, P, c_int, c_int)
_sig("ct_trange_d", P, P, P, P, c_int)
_sig("ct_stochastic_d", P, P, P, P, P, c_int, c_int, c_int)
_sig("ct_aroon_d", P, P, P, P, c_int, c_int)
_sig("ct_aroonosc_d", P, P, P, c_int, c_int)
_sig("ct_plus_dm_d", P, P, P, c_int, c_int)
_sig("ct_minus_dm_d", P, P, P, c_int, c_int)
# dense batch on device (include/tacuda.h)
for _n in ("sma", "ema", "rsi", "wma", "dema", "tema", "stddev", "var", "sum", "max", "min"):
    _sig(f"ct_{_n}_batch_d", P, P, c_int, c_int, c_int, c_int)
for _n in ("natr", "adx", "dx", "plus_di", "minus_di", "willr", "cci"):
    _sig(f"ct_{_n}_batch_d", P, P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_atr_batch_d", P, P, P, P, c_int, c_int, c_int, c_float, c_int)
_sig("ct_macd_batch_d", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int)
_sig("ct_bbands_batch_d", P, P, P, P, c_int, c_int, c_int, c_float, c_float, c_int)
_sig("ct_stochastic_batch_d", P, P, P, P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ct_kama_batch_d", P, P, c_int, c_int, c_int, c_int, c_int, c_int)
_sig("ct_t3_batch_d", P, P, c_int, c_int, c_int, c_float, c_int)
_sig("ct_trange_batch_d", P, P, P, P, c_int, c_int, c_int)
_sig("ct_aroonosc_batch_d", P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_aroon_batch_d", P, P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_plus_dm_batch_d", P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_minus_dm_batch_d", P, P, P, c_int, c_int, c_int, c_int)
# host calls (include/tacuda.h): host arrays in and out, the call waits
for _n in ("sma", "ema", "wma", "dema", "tema", "rsi", "stddev", "var", "sum", "max", "min"):
    _sig(f"ct_{_n}", P, P, c_int, c_int)
    _sig(f"ct_{_n}_batch", P, P, c_int, c_int, c_int, c_int)
_sig("ct_t3", P, P, c_int, c_int, c_float)
_sig("ct_kama", P, P, c_int, c_int, c_int, c_int)
_sig("ct_ma", P, P, c_int, c_int, c_int)
_sig("ct_macd", P, P, P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ct_bbands", P, P, P, P, c_int, c_int, c_float, c_float)
_sig("ct_atr", P, P, P, P, c_int, c_int, c_float)
for _n in ("natr", "adx", "dx", "plus_di", "minus_di", "willr", "cci"):
    _sig(f"ct_{_n}", P, P, P, P, c_int, c_int)
    _sig(f"ct_{_n}_batch", P, P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_trange", P, P, P, P, c_int)
_sig("ct_stochastic", P, P, P, P, P, c_int, c_int, c_int)
_sig("ct_plus_dm", P, P, P, c_int, c_int)
_sig("ct_minus_dm", P, P, P, c_int, c_int)
_sig("ct_aroon", P, P, P, P, P, c_int, c_int, c_int)
_sig("ct_aroonosc", P, P, P, c_int, c_int)
_sig("ct_atr_batch", P, P, P, P, c_int, c_int, c_int, c_float, c_int)
_sig("ct_macd_batch", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int)
_sig("ct_bbands_batch", P, P, P, P, c_int, c_int, c_int, c_float, c_float, c_int)
_sig("ct_stochastic_batch", P, P, P, P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ct_kama_batch", P, P, c_int, c_int, c_int, c_int, c_int, c_int)
_sig("ct_t3_batch", P, P, c_int, c_int, c_int, c_float, c_int)
_sig("ct_trange_batch", P, P, P, P, c_int, c_int, c_int)
_sig("ct_aroonosc_batch", P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_aroon_batch", P, P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_plus_dm_batch", P, P, P, c_int, c_int, c_int, c_int)
_sig("ct_minus_dm_batch", P, P, P, c_int, c_int, c_int, c_int)
# device aliases and the moving-average selector (include/tacuda.h)
for _n in ("sma", "ema", "wma", "dema", "tema", "rsi", "stddev", "var", "sum", "max", "min"):
    _sig(f"ct_{_n}_device", P, P, c_int, c_int)
_sig("ct_t3_device", P, P, c_int, c_int, c_float)
_sig("ct_kama_device", P, P, c_int, c_int, c_int, c_int)
_sig("ct_ma_device", P, P, c_int, c_int, c_int)
_sig("ct_ma_d", P, P, c_int, c_int, c_int)
_sig("ct_ma_batch", P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ct_ma_batch_d", P, P, c_int, c_int, c_int, c_int, c_int)
# typed variants: storage type, compute type (include/tacuda.h, tacuda_device.h)
for _n in ("sma", "ema", "rsi"):
    _sig(f"ct_{_n}_ex", P, P, c_int, c_int, c_int, c_int)
    _sig(f"ct_{_n}_ex_d", P, P, c_int, c_int, c_int, c_int)
for _sfx in ("_ex", "_ex_d"):
    _sig(f"ct_atr{_sfx}", P, P, P, P, c_int, c_int, c_float, c_int, c_int)
    _sig(f"ct_macd{_sfx}", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int)
    _sig(f"ct_bbands{_sfx}", P, P, P, P, c_int, c_int, c_float, c_float, c_int, c_int)
# TRIMA and MACDEXT (a moving-average type per line)
_sig("ct_trima_d", P, P, c_int, c_int)
_sig("ct_trima_device", P, P, c_int, c_int)
_sig("ct_trima", P, P, c_int, c_int)
_sig("ct_trima_batch", P, P, c_int, c_int, c_int, c_int)
_sig("ct_trima_batch_d", P, P, c_int, c_int, c_int, c_int)
_sig("ct_macdext_d", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int)
_sig("ct_macdext", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int)
for _sfx in ("_batch", "_batch_d"):
    _sig(f"ct_macdext{_sfx}", P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
         c_int)
_sig("ct_device_upload_ex", P, P, c_int, c_int)
_sig("ct_device_download_ex", P, P, c_int, c_int)
_sig("ctx_host_staging", c_int, stream=False)
_sig("ct_device_buffer_pool_cleanup", stream=False, restype=None)
_sig("ct_device_buffer_pool_allocation_count", stream=False, restype=ctypes.c_size_t)
_sig("ct_device_buffer_pool_set_max_cache_bytes", ctypes.c_size_t, stream=False,
     restype=ctypes.c_size_t)
# memory / errors
_sig("ct_device_alloc", c_int, stream=False, restype=c_void_p)
_sig("ct_device_free", c_void_p, stream=False, restype=None)
_sig("ct_device_upload", P, P, c_int)
_sig("ct_device_download", P, P, c_int)
_sig("ct_device_sync")
_sig("ct_get_last_error", stream=False, restype=ctypes.c_char_p)
_sig("ct_status_string", c_int, stream=False, restype=ctypes.c_char_p)
_sig("ct_version", stream=False, restype=ctypes.c_char_p)
# extensions (include/tacuda_amd.h)
_sig("ctx_synth_ohlcv", ctypes.c_ulonglong, c_int, c_int, c_int, P, P, P, P, P, c_int, stream=False)
_sig("ctx_workspace_reserve", c_void_p, ctypes.c_size_t, stream=False)
_sig("ctx_kernel_timing", c_int, stream=False)
_sig("ctx_kernel_times", ctypes.c_char_p, c_int, ctypes.POINTER(c_float), c_int,
     ctypes.POINTER(c_int), stream=False)


class PackParams(ctypes.Structure):
    """ctxPackParams of include/tacuda_amd.h."""
    _fields_ = [(n, c_int) for n in ("sma_period", "ema_period", "wma_period", "dema_period",
                                      "tema_period", "rsi_period", "macd_fast", "macd_slow",
                                      "macd_signal", "bb_period")] + \
        [("bb_up", c_float), ("bb_dn", c_float)] + \
        [(n, c_int) for n in ("atr_period", "adx_period", "stoch_k", "stoch_d", "willr_period",
                              "cci_period")]


PACK_OUTPUTS = ("sma", "ema", "wma", "dema", "tema", "rsi", "macd", "macd_signal", "macd_hist",
                "bb_upper", "bb_middle", "bb_lower", "atr", "adx", "stoch_k", "stoch_d",
                "willr", "cci")


class PackOutputs(ctypes.Structure):
    """ctxPackOutputs of include/tacuda_amd.h (device pointers, NULL = skip)."""
    _fields_ = [(n, c_void_p) for n in PACK_OUTPUTS]


_sig("ctx_pack_batch_d", P, P, P, c_int, c_int, ctypes.POINTER(PackParams),
     ctypes.POINTER(PackOutputs))
# ragged layouts (include/tacuda_amd.h)
_sig("ctx_ragged_create", P, c_int, ctypes.POINTER(c_void_p), stream=False)
_sig("ctx_ragged_destroy", c_void_p, stream=False)
_sig("ctx_pack_ragged_d", P, P, P, c_void_p, ctypes.POINTER(PackParams),
     ctypes.POINTER(PackOutputs))


class ScanParams(ctypes.Structure):
    """ctxScanParams of include/tacuda_amd.h."""
    _fields_ = [("ema_period", c_int), ("kama_period", c_int), ("kama_fast", c_int),
                ("kama_slow", c_int), ("t3_period", c_int), ("t3_vfactor", c_float),
                ("sma_period", c_int)]


SCAN_OUTPUTS = ("ema", "kama", "t3", "sma")


class ScanOutputs(ctypes.Structure):
    """ctxScanOutputs of include/tacuda_amd.h (device pointers, NULL = skip)."""
    _fields_ = [(n, c_void_p) for n in SCAN_OUTPUTS]


_sig("ctx_scan_batch_d", P, c_int, c_int, ctypes.POINTER(ScanParams), ctypes.POINTER(ScanOutputs))
for _n in ("sma", "ema", "wma", "dema", "tema", "rsi", "var", "stddev", "sum", "max", "min"):
    _sig(f"ctx_{_n}_ragged_d", P, P, c_void_p, c_int)
for _n in ("atr", "natr", "adx", "dx", "plus_di", "minus_di", "willr", "cci"):
    _sig(f"ctx_{_n}_ragged_d", P, P, P, P, c_void_p, c_int)
_sig("ctx_t3_ragged_d", P, P, c_void_p, c_int, c_float)
_sig("ctx_kama_ragged_d", P, P, c_void_p, c_int, c_int, c_int)
_sig("ctx_macd_ragged_d", P, P, P, P, c_void_p, c_int, c_int, c_int)
_sig("ctx_bbands_ragged_d", P, P, P, P, c_void_p, c_int, c_float, c_float)
_sig("ctx_stochastic_ragged_d", P, P, P, P, P, c_void_p, c_int, c_int)
_sig("ctx_trange_ragged_d", P, P, P, P, c_void_p)
# persistent tuned plans (include/tacuda_amd.h)
_sig("ctx_pack_plan_create", P, P, P, c_int, c_int, c_void_p, ctypes.POINTER(PackParams),
     ctypes.POINTER(PackOutputs), c_int, ctypes.POINTER(c_void_p), stream=False)
_sig("ctx_plan_launch", c_void_p, c_void_p, stream=False)
_sig("ctx_plan_info", c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
     ctypes.POINTER(c_float), ctypes.POINTER(c_float), stream=False)
_sig("ctx_plan_destroy", c_void_p, stream=False)
# streaming (include/tacuda_stream.h)
_PP = ctypes.POINTER(c_void_p)
for _n in ("sma", "ema", "wma", "rsi", "atr", "natr"):
    _sig(f"ct_{_n}_stream_create", c_int, _PP, stream=False)
    _sig(f"ct_{_n}_stream_batch_create", c_int, c_int, _PP, stream=False)
_sig("ct_obv_stream_create", _PP, stream=False)
_sig("ct_ad_stream_create", _PP, stream=False)
_sig("ct_macd_stream_create", c_int, c_int, c_int, _PP, stream=False)
_sig("ct_bbands_stream_create", c_int, c_float, c_float, _PP, stream=False)
_sig("ct_obv_stream_batch_create", c_int, _PP, stream=False)
_sig("ct_ad_stream_batch_create", c_int, _PP, stream=False)
_sig("ct_macd_stream_batch_create", c_int, c_int, c_int, c_int, _PP, stream=False)
_sig("ct_bbands_stream_batch_create", c_int, c_float, c_float, c_int, _PP, stream=False)
for _pre in ("ct_{}_stream_append", "ct_{}_stream_batch_append"):
    for _n in ("sma", "ema", "wma", "rsi"):
        _sig(_pre.format(_n), c_void_p, P, c_int, P, stream=False)
    for _n in ("atr", "natr"):
        _sig(_pre.format(_n), c_void_p, P, P, P, c_int, P, stream=False)
    _sig(_pre.format("obv"), c_void_p, P, P, c_int, P, stream=False)
    _sig(_pre.format("ad"), c_void_p, P, P, P, P, c_int, P, stream=False)
    _sig(_pre.format("macd"), c_void_p, P, c_int, P, P, P, stream=False)
    _sig(_pre.format("bbands"), c_void_p, P, c_int, P, P, P, stream=False)
for _n in ("destroy", "reset"):
    _sig(f"ct_stream_{_n}", c_void_p, stream=False)
    _sig(f"ct_stream_batch_{_n}", c_void_p, stream=False)
_sig("ct_stream_lookback", c_void_p, ctypes.POINTER(c_int), stream=False)
_sig("ct_stream_outputs", c_void_p, ctypes.POINTER(c_int), stream=False)
_sig("ct_stream_bars", c_void_p, ctypes.POINTER(ctypes.c_longlong), stream=False)
_sig("ct_stream_batch_lookback", c_void_p, ctypes.POINTER(c_int), stream=False)
_sig("ct_stream_batch_outputs", c_void_p, ctypes.POINTER(c_int), stream=False)
_sig("ct_stream_batch_series", c_void_p, ctypes.POINTER(c_int), stream=False)
_sig("ct_stream_batch_bars", c_void_p, ctypes.POINTER(ctypes.c_longlong), stream=False)
_sig("ct_stream_batch_set_exec", c_void_p, c_int, stream=False)
_sig("ctx_stream_batch_append_d", c_void_p, P, P, P, P, c_int, P, P, P)
# candlestick patterns (include/tacuda_device.h, include/tacuda_amd.h)
_sig("ctx_cdl_count", stream=False)
_sig("ctx_cdl_name", c_int, stream=False, restype=ctypes.c_char_p)
_sig("ctx_cdl_batch_d", c_int, P, P, P, P, P, c_int, c_int)
_sig("ctx_cdl_all_batch_d", P, P, P, P, P, c_int, c_int)


def _cdl_sigs(names):
    for n in names:
        _sig(f"ct_cdl_{n}_d", P, P, P, P, P, c_int)
        _sig(f"ct_cdl_{n}", P, P, P, P, P, c_int)
        _sig(f"ct_cdl_{n}_batch_d", P, P, P, P, P, c_int, c_int, c_int)
        _sig(f"ct_cdl_{n}_batch", P, P, P, P, P, c_int, c_int, c_int)


_cdl_sigs(("bullish_engulfing", "bearish_engulfing"))
# differentiable indicators (include/tacuda_amd.h)
_D = ctypes.c_double
for _n in ("ema", "rsi"):
    _sig(f"ctx_{_n}_alpha_fwd_d", P, P, P, c_int, c_int, c_int, _D)
    _sig(f"ctx_{_n}_alpha_bwd_d", P, P, P, P, P, c_int, c_int, c_int, _D)
    _sig(f"ctx_{_n}_alpha_fwd_dp", P, P, P, c_int, c_int, c_int, P)
    _sig(f"ctx_{_n}_alpha_bwd_dp", P, P, P, P, P, c_int, c_int, c_int, P)
_sig("ctx_alpha_ckpt_per_row", stream=False)
# price transforms, momentum, window midpoints, linear regression
for _n in ("momentum", "roc", "rocp", "rocr", "rocr100", "midpoint", "linearreg",
           "linearreg_slope", "linearreg_intercept", "linearreg_angle", "tsf"):
    _sig(f"ct_{_n}_d", P, P, c_int, c_int)
    _sig(f"ct_{_n}", P, P, c_int, c_int)
    _sig(f"ct_{_n}_batch_d", P, P, c_int, c_int, c_int, c_int)
    _sig(f"ct_{_n}_batch", P, P, c_int, c_int, c_int, c_int)
for _n in ("momentum", "roc", "rocp", "rocr", "rocr100"):
    _sig(f"ct_{_n}_device", P, P, c_int, c_int)
for _n, _k in (("avgprice", 4), ("medprice", 2), ("typprice", 3), ("wclprice", 3)):
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, *([P] * (_k + 1)), c_int)
    for _f in (f"ct_{_n}_batch_d", f"ct_{_n}_batch"):
        _sig(_f, *([P] * (_k + 1)), c_int, c_int, c_int)
for _f in ("ct_midprice_d", "ct_midprice"):
    _sig(_f, P, P, P, c_int, c_int)
for _f in ("ct_midprice_batch_d", "ct_midprice_batch"):
    _sig(_f, P, P, P, c_int, c_int, c_int, c_int)
# element-wise math transforms and operators; BOP, IBS, CHANGE, AVGDEV, MINMAX
MATH_UNARY = ("acos", "asin", "atan", "ceil", "cos", "cosh", "exp", "floor", "ln", "log10",
              "sin", "sinh", "sqrt", "tan", "tanh")
MATH_BINARY = ("add", "sub", "mult", "div")
for _n in MATH_UNARY:
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, P, P, c_int)
for _n in MATH_BINARY:
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, P, P, P, c_int)
for _n in ("change", "avgdev", "rolling_mean", "rolling_max", "rolling_min", "rolling_std",
           "pivot_high", "pivot_low", "skew", "trendscore", "market_strength"):
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, P, P, c_int, c_int)
    for _f in (f"ct_{_n}_batch_d", f"ct_{_n}_batch"):
        _sig(_f, P, P, c_int, c_int, c_int, c_int)
_sig("ct_change_device", P, P, c_int, c_int)
for _n, _k in (("bop", 4), ("ibs", 3)):
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, *([P] * (_k + 1)), c_int)
    for _f in (f"ct_{_n}_batch_d", f"ct_{_n}_batch"):
        _sig(_f, *([P] * (_k + 1)), c_int, c_int, c_int)
for _f in ("ct_minmax_d", "ct_minmax"):
    _sig(_f, P, P, P, c_int, c_int)
for _f in ("ct_minmax_batch_d", "ct_minmax_batch"):
    _sig(_f, P, P, P, c_int, c_int, c_int, c_int)
# TA-Lib 0.8.3 row functions: name -> (inputs, outputs, periods)
ROW_FUNCTIONS = {"cumsum": (1, 1, 0), "pvt": (2, 1, 0), "rma": (1, 1, 1), "zlema": (1, 1, 1),
                 "efi": (2, 1, 1), "vwma": (2, 1, 1), "er": (1, 1, 1), "vhf": (1, 1, 1),
                 "dpo": (1, 1, 1), "chop": (3, 1, 1), "adr": (2, 1, 1), "ao": (2, 1, 2),
                 "qstick": (2, 1, 1), "donchian": (2, 3, 1), "marketfi": (3, 1, 0),
                 "wad": (3, 1, 0), "mcgd": (1, 1, 1), "rvol": (1, 1, 1), "cmf": (4, 1, 1),
                 "emv": (3, 1, "1d"), "fosc": (1, 1, 1), "eri": (3, 2, 1),
                 "percentrank": (1, 1, 1), "si": (4, 1, "d"), "asi_wilder": (4, 1, "d"),
                 "vortex": (3, 2, 1), "cmou": (1, 1, 1)}
_PARAMS = {0: [], 1: [c_int], 2: [c_int, c_int], "d": [ctypes.c_double],
           "1d": [c_int, ctypes.c_double]}
for _n, (_i, _o, _p) in ROW_FUNCTIONS.items():
    for _f in (f"ct_{_n}_d", f"ct_{_n}"):
        _sig(_f, *([P] * (_i + _o)), c_int, *_PARAMS[_p])
    for _f in (f"ct_{_n}_batch_d", f"ct_{_n}_batch"):
        _sig(_f, *([P] * (_i + _o)), c_int, c_int, *_PARAMS[_p], c_int)
# OBV and AD (running sums on the look-back scan engine)
for _f in ("ct_obv_d", "ct_obv"):
    _sig(_f, P, P, P, c_int)
for _f in ("ct_obv_batch_d", "ct_obv_batch"):
    _sig(_f, P, P, P, c_int, c_int, c_int)
for _f in ("ct_ad_d", "ct_ad"):
    _sig(_f, P, P, P, P, P, c_int)
for _f in ("ct_ad_batch_d", "ct_ad_batch"):
    _sig(_f, P, P, P, P, P, c_int, c_int, c_int)
# ULTOSC and MFI
for _f in ("ct_ultosc_d", "ct_ultosc"):
    _sig(_f, P, P, P, P, c_int, c_int, c_int, c_int)
for _f in ("ct_ultosc_batch_d", "ct_ultosc_batch"):
    _sig(_f, P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int)
for _f in ("ct_mfi_d", "ct_mfi"):
    _sig(_f, P, P, P, P, P, c_int, c_int)
for _f in ("ct_mfi_batch_d", "ct_mfi_batch"):
    _sig(_f, P, P, P, P, P, c_int, c_int, c_int, c_int)
_sig("ctx_macd_bwd_d", P, P, P, P, c_int, c_int, c_int, c_int, c_int)
_sig("ctx_bbands_bwd_d", P, P, P, P, P, c_int, c_int, c_int, c_float, c_float)


class TacudaError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


def lib() -> ctypes.CDLL:
    """Load the HIP library; raise if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(tacuda_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        fn = L.ctx_cdl_name
        fn.restype = ctypes.c_char_p
        fn.argtypes = [c_int]
        _cdl_sigs([L.ctx_cdl_name(k).decode() for k in range(L.ctx_cdl_count())])
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error() -> str:
    return lib().ct_get_last_error().decode()


def call(name: str, *args) -> int:
    """Call an entry point and raise TacudaError on a non-success status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise TacudaError(name, rc, last_error())
    return rc


def synth_ohlcv(n: int, rows: int = 1, seed: int = 0, first: int = 0, threads: int = 0):
    """Seeded float32 OHLCV as numpy arrays of shape [rows, n] (or [n])."""
    import numpy as np

    arrs = [np.empty((rows, n), dtype=np.float32) for _ in range(5)]
    call("ctx_synth_ohlcv", seed, first, rows, n, *[a.ctypes.data for a in arrs], threads)
    if rows == 1:
        arrs = [a[0] for a in arrs]
    return dict(zip(("open", "high", "low", "close", "volume"), arrs))


def kernel_timing(enable: bool) -> None:
    """Record HIP events around the fused-pass, long-scan and candle kernels
    launched from now on (ctx_kernel_timing)."""
    call("ctx_kernel_timing", 1 if enable else 0)


def kernel_times(max_n: int = 4096):
    """[(kernel name, milliseconds)] of the launches recorded since the last
    call, in launch order (ctx_kernel_times waits for them)."""
    names = ctypes.create_string_buffer(max_n * 32)
    ms = (c_float * max_n)()
    n = c_int(0)
    call("ctx_kernel_times", names, len(names), ms, max_n, ctypes.byref(n))
    keys = names.value.decode().split("\n") if n.value else []
    return list(zip(keys, [ms[i] for i in range(n.value)]))


from . import ops  # noqa: E402,F401
from . import stream  # noqa: E402,F401
from . import autograd  # noqa: E402,F401
