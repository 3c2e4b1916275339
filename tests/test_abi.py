"""CPU checks of the drop-in boundary: the HIP library loads without a GPU
and exports every entry point declared in include/*.h; error plumbing and
the synthetic-data generator work without touching the device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        # drop macro definitions, continuation lines included
        text = re.sub(r"^\s*#define[^\n]*(?:\\\n[^\n]*)*", "", text, flags=re.M)
        # plain declarations
        for m in re.finditer(r"\b(ct_\w+|ctx_\w+)\s*\(", text):
            names.add(m.group(1))
        # batch macros: CT_BATCH1(name, ...) and CT_BATCH_HLC(name)
        for m in re.finditer(r"CT_BATCH1\((\w+),", text):
            names.update({f"ct_{m.group(1)}_batch", f"ct_{m.group(1)}_batch_d"})
        for m in re.finditer(r"CT_BATCH_HLC\((\w+)\)", text):
            names.update({f"ct_{m.group(1)}_batch", f"ct_{m.group(1)}_batch_d"})
        for m in re.finditer(r"CT_CDL_DECL\((\w+)\)", text):
            names.update({f"ct_cdl_{m.group(1)}_d", f"ct_cdl_{m.group(1)}_batch_d",
                          f"ct_cdl_{m.group(1)}", f"ct_cdl_{m.group(1)}_batch"})
        # ragged macros: CTX_RAGGED1(name) and CTX_RAGGED_HLC(name)
        for m in re.finditer(r"CTX_RAGGED(?:1|_HLC)\((\w+)\)", text):
            names.add(f"ctx_{m.group(1)}_ragged_d")
    names.discard("ct_")
    return sorted(n for n in names if not n.endswith("_") and "##" not in n)


def _lib_path():
    return os.path.join(ROOT, "tacuda_amd", "libtacuda_amd.so")


def test_library_loads_and_exports_every_declared_symbol():
    path = _lib_path()
    assert os.path.exists(path), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(path)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert len(_declared()) > 100


def test_only_declared_symbols_are_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    ours = {n for n in exported if n.startswith(("ct_", "ctx_"))}
    assert ours == set(_declared()) & exported
    assert ours >= set(_declared())


def test_status_strings_and_version():
    import tacuda_amd as ta

    L = ta.lib()
    assert L.ct_version() == b"0.9.3"
    assert L.ct_status_string(4) == b"invalid argument"
    assert L.ct_status_string(0) == b"success"


def test_invalid_arguments_fail_before_touching_the_device():
    import tacuda_amd as ta

    L = ta.lib()
    assert L.ct_ema_d(None, None, 10, 5, None) == 4
    assert b"null" in L.ct_get_last_error()
    buf = np.zeros(4, np.float32)
    p = buf.ctypes.data
    assert L.ct_sma_d(p, p, 0, 5, None) == 4
    assert L.ct_sma_d(p, p, 4, 0, None) == 4
    assert L.ct_t3_d(p, p, 4, 1, 0.7, None) == 4
    assert L.ct_stochastic_d(p, p, p, p, p, 4, 14, 0, None) == 4
    # moving-average types outside the hot path, and storage types other
    # than float32, are refused before any device work
    assert L.ct_ma_device(p, p, 4, 3, 7, None) == 6  # MAMA
    assert L.ct_ma_device(p, p, 4, 3, 99, None) == 6
    assert L.ct_sma_ex_d(p, p, 4, 3, 2, 2, None) == 6  # float64 storage
    assert L.ct_rsi_ex_d(p, p, 4, 3, 0, 2, None) == 6  # half storage needs float32 compute
    assert L.ct_ema_ex_d(p, p, 4, 3, 7, 1, None) == 6  # not a data type
    assert L.ct_sma_ex(None, p, 4, 3, 1, 1, None) == 4
    assert L.ctx_host_staging(3) == 4
    # MACDEXT: MAMA and the moving averages above T3 are not provided; TA-Lib's
    # period ranges are enforced
    assert L.ct_macdext_d(p, p, p, p, 4, 2, 7, 3, 0, 2, 0, None) == 6
    assert L.ct_macdext_d(p, p, p, p, 4, 2, 0, 3, 12, 2, 0, None) == 6
    assert L.ct_macdext_d(p, p, p, p, 4, 2, 0, 3, 0, 2, 99, None) == 4
    assert L.ct_macdext_d(p, p, p, p, 4, 1, 0, 3, 0, 2, 0, None) == 4
    assert L.ct_macd_d(p, p, p, p, 4, 12, 26, 9, 7, None) == 6
    assert L.ct_trima_d(p, p, 4, 0, None) == 4


def test_host_aroon_short_series_is_all_nan_without_the_device():
    """Host AROON with a period beyond the series: TA-Lib's zero-count
    convention, every line NaN (reference Aroon.cu:60-65), and invalid
    periods are rejected."""
    import tacuda_amd as ta

    L = ta.lib()
    h = np.arange(5, dtype=np.float32) + 10
    l = h - 1
    up, dn, osc = (np.zeros(5, np.float32) for _ in range(3))
    args = [h.ctypes.data, l.ctypes.data, up.ctypes.data, dn.ctypes.data, osc.ctypes.data, 5]
    assert L.ct_aroon(*args, 3, 6, None) == 0
    assert np.isnan(up).all() and np.isnan(dn).all() and np.isnan(osc).all()
    assert L.ct_aroon(*args, 0, 3, None) == 4


def test_synthetic_generator_is_seeded_and_well_formed():
    import tacuda_amd as ta

    a = ta.synth_ohlcv(5000, rows=3, seed=0)
    b = ta.synth_ohlcv(5000, rows=3, seed=0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    one = ta.synth_ohlcv(5000, rows=1, seed=0, first=2)
    assert np.array_equal(one["close"], a["close"][2])
    assert not np.array_equal(a["close"][0], a["close"][1])
    o, h, l, c = a["open"], a["high"], a["low"], a["close"]
    assert (c[:, 0] == 100).all() and (o[:, 0] == 100).all()
    assert np.array_equal(o[:, 1:], c[:, :-1])
    assert (h >= np.maximum(o, c)).all() and (l <= np.minimum(o, c)).all()
    r = np.diff(np.log(c.astype(np.float64)), axis=1)
    assert abs(r.std() - 0.01) < 0.001 and abs(r.mean()) < 0.001
    v = np.log(a["volume"].astype(np.float64))
    assert abs(v.mean() - 10) < 0.1 and abs(v.std() - 1) < 0.1


def test_pack_arguments_are_validated_without_the_device():
    import tacuda_amd as ta
    from tacuda_amd.ops import PACK_DEFAULTS

    L = ta.lib()
    buf = np.zeros(8, np.float32)
    p = buf.ctypes.data
    prm = ta.PackParams(**PACK_DEFAULTS)
    none = ta.PackOutputs()
    assert L.ctx_pack_batch_d(p, p, p, 8, 1, prm, none, None) == 4
    assert b"no output" in L.ct_get_last_error()
    assert L.ctx_pack_batch_d(p, p, p, 0, 1, prm, ta.PackOutputs(ema=p), None) == 4
    # EMA, DEMA and TEMA share one cascade in a pack
    mixed = ta.PackParams(**{**PACK_DEFAULTS, "dema_period": 30})
    assert L.ctx_pack_batch_d(p, p, p, 8, 1, mixed, ta.PackOutputs(ema=p, dema=p), None) == 6
    bad = ta.PackParams(**{**PACK_DEFAULTS, "rsi_period": 0})
    assert L.ctx_pack_batch_d(p, p, p, 8, 1, bad, ta.PackOutputs(rsi=p), None) == 4
    # STOCH smoothing of any length, up to the pack's 1024-bar window halo
    deep = ta.PackParams(**{**PACK_DEFAULTS, "stoch_d": 510})
    assert L.ctx_pack_batch_d(p, p, p, 8, 1, deep, ta.PackOutputs(stoch_k=p), None) == 6
    wide = ta.PackParams(**{**PACK_DEFAULTS, "sma_period": 1026})
    assert L.ctx_pack_batch_d(p, p, p, 8, 1, wide, ta.PackOutputs(sma=p), None) == 6
    # high/low are required only by the outputs that read them
    assert L.ctx_pack_batch_d(None, None, p, 8, 1, prm, ta.PackOutputs(atr=p), None) == 4


def test_scan_arguments_are_validated_without_the_device():
    import tacuda_amd as ta
    from tacuda_amd.ops import SCAN_DEFAULTS

    L = ta.lib()
    buf = np.zeros(8, np.float32)
    p = buf.ctypes.data
    prm = ta.ScanParams(**SCAN_DEFAULTS)
    assert L.ctx_scan_batch_d(p, 8, 1, prm, ta.ScanOutputs(), None) == 4
    assert b"no output" in L.ct_get_last_error()
    assert L.ctx_scan_batch_d(p, 0, 1, prm, ta.ScanOutputs(ema=p), None) == 4
    assert L.ctx_scan_batch_d(None, 8, 1, prm, ta.ScanOutputs(ema=p), None) == 4
    t3 = ta.ScanParams(**{**SCAN_DEFAULTS, "t3_period": 1})
    assert L.ctx_scan_batch_d(p, 8, 1, t3, ta.ScanOutputs(t3=p), None) == 4
    bad = ta.ScanParams(**{**SCAN_DEFAULTS, "kama_fast": 0})
    assert L.ctx_scan_batch_d(p, 8, 1, bad, ta.ScanOutputs(kama=p), None) == 4
    wide = ta.ScanParams(**{**SCAN_DEFAULTS, "sma_period": 1026})
    assert L.ctx_scan_batch_d(p, 8, 1, wide, ta.ScanOutputs(sma=p), None) == 6
    wide = ta.ScanParams(**{**SCAN_DEFAULTS, "kama_period": 1024})
    assert L.ctx_scan_batch_d(p, 8, 1, wide, ta.ScanOutputs(kama=p), None) == 6


def test_ragged_layout_arguments_are_validated_without_the_device():
    import ctypes

    import tacuda_amd as ta

    L = ta.lib()
    h = ctypes.c_void_p()
    off = np.array([0, 10, 5], dtype=np.int64)
    assert L.ctx_ragged_create(off.ctypes.data, 2, ctypes.byref(h)) == 4
    assert b"decrease" in L.ct_get_last_error()
    assert L.ctx_ragged_create(off.ctypes.data, 0, ctypes.byref(h)) == 4
    assert L.ctx_ragged_create(None, 2, ctypes.byref(h)) == 4
    empty = np.array([0, 0, 0], dtype=np.int64)
    assert L.ctx_ragged_create(empty.ctypes.data, 2, ctypes.byref(h)) == 4
    assert L.ctx_ragged_destroy(None) == 0
    buf = np.zeros(4, np.float32)
    assert L.ctx_ema_ragged_d(buf.ctypes.data, buf.ctypes.data, None, 5, None) == 4


def test_kernel_timing_without_the_device():
    import tacuda_amd as ta

    L = ta.lib()
    n = ctypes.c_int(-1)
    assert L.ctx_kernel_times(None, 0, None, 0, None) == 4
    assert L.ctx_kernel_times(None, 0, None, 1, ctypes.byref(n)) == 4
    assert L.ctx_kernel_times(None, 0, None, 0, ctypes.byref(n)) == 0 and n.value == 0
    assert L.ctx_kernel_timing(1) == 0 and L.ctx_kernel_timing(0) == 0
    assert ta.kernel_times() == []


def test_fault_hook_only_in_the_fault_injection_build():
    """ct_test_stream_batch_fail_next exists only in the test build
    libtacuda_amd_faults.so, not in the product library."""
    prod = subprocess.run(["nm", "-D", "--defined-only", _lib_path()], capture_output=True,
                          text=True, check=True).stdout
    assert "ct_test_stream_batch_fail_next" not in prod
    faults = os.path.join(ROOT, "tacuda_amd", "libtacuda_amd_faults.so")
    test = subprocess.run(["nm", "-D", "--defined-only", faults], capture_output=True, text=True,
                          check=True).stdout
    assert "ct_test_stream_batch_fail_next" in test


def test_reference_abi_of_the_hot_path():
    """Every hot-path entry point the reference's public headers declare
    (tests/golden/reference_abi.json, written by make_abi_fixture.py from
    include/tacuda.h, tacuda_device.h and tacuda_stream.h: the indicator
    families' host, device, batch and typed forms, the 61 candle calls and
    the legacy engulfing names, the streaming surface, memory and error
    calls) is declared here with the same return and argument types
    (hipStream_t for cudaStream_t) and exported by the library."""
    import json
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_abi_fixture import prototypes

    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_abi.json")))
    assert len(fx) > 400
    ours = {}
    for h in HEADERS:
        txt = subprocess.run(["gcc", "-E", "-P", "-x", "c", h], capture_output=True, text=True,
                             check=True).stdout
        ours.update(prototypes(txt.replace('__attribute__((visibility("default")))',
                                           "CTAPI_EXPORT")))
    missing = sorted(n for n in fx if n not in ours)
    assert not missing, f"not declared: {missing}"
    differ = {n: (fx[n], ours[n]) for n in fx if fx[n] != ours[n]}
    assert not differ, f"prototypes differ: {differ}"
    lib = ctypes.CDLL(_lib_path())
    unexported = [n for n in fx if not hasattr(lib, n)]
    assert not unexported, f"not exported: {unexported}"
