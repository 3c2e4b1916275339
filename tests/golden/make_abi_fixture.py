"""Writes tests/golden/reference_abi.json: the reference's C prototypes of
the hot-path entry points (name -> argument type list, cudaStream_t written
as hipStream_t), read from its public headers include/tacuda.h,
include/tacuda_device.h and include/tacuda_stream.h.  Run in a container
that has the reference checkout; the JSON is committed and the ABI test
reads only it."""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_abi.json")

# indicator families the build provides (every host, device, batch and
# typed form the reference declares for them is taken)
FAMILIES = ["sma", "ema", "wma", "dema", "tema", "t3", "kama", "trima", "rsi", "macd",
            "macdext", "bbands", "atr", "natr", "trange", "adx", "dx", "plus_di", "minus_di",
            "plus_dm", "minus_dm", "stochastic", "willr", "cci", "aroon", "aroonosc", "stddev",
            "var", "sum", "min", "max", "ma",
            # price transforms, momentum, window midpoints, linear regression
            "avgprice", "medprice", "typprice", "wclprice", "momentum", "roc", "rocp", "rocr",
            "rocr100", "midprice", "midpoint", "linearreg", "linearreg_slope",
            "linearreg_intercept", "linearreg_angle", "tsf",
            # element-wise math, bar and window statistics
            "acos", "asin", "atan", "ceil", "cos", "cosh", "exp", "floor", "ln", "log10", "sin",
            "sinh", "sqrt", "tan", "tanh", "add", "sub", "mult", "div", "bop", "ibs", "change",
            "avgdev", "minmax",
            # TA-Lib 0.8.3 row functions
            "cumsum", "pvt", "rma", "zlema", "efi", "vwma", "er", "vhf", "dpo", "chop", "adr",
            "ao", "qstick", "donchian", "marketfi", "wad", "mcgd", "rvol", "cmf", "emv", "fosc",
            "eri", "percentrank", "si", "asi_wilder", "obv", "ad", "ultosc", "mfi",
            "vortex", "cmou", "rolling_mean", "rolling_max", "rolling_min", "rolling_std",
            "pivot_high", "pivot_low", "skew", "trendscore", "market_strength"]
SUFFIXES = ["", "_d", "_device", "_batch", "_batch_d", "_ex", "_ex_d"]
EXTRA = ["ct_status_string", "ct_get_last_error", "ct_version", "ct_device_alloc",
         "ct_device_alloc_ex", "ct_device_free", "ct_device_free_ex", "ct_device_upload",
         "ct_device_upload_ex", "ct_device_download", "ct_device_download_ex", "ct_device_sync",
         "ct_device_buffer_pool_cleanup", "ct_device_buffer_pool_allocation_count",
         "ct_device_buffer_pool_set_max_cache_bytes"]


def norm_type(t):
    t = re.sub(r"\bcudaStream_t\b", "hipStream_t", t)
    t = re.sub(r"\s+", " ", t).strip()
    t = re.sub(r"\s*\*\s*", "*", t)
    # drop the parameter name: the last identifier unless the type is bare
    m = re.match(r"^(.*?[\w*])\s*\b([A-Za-z_]\w*)$", t)
    if m and m.group(1).strip() not in ("const", "struct", "unsigned"):
        t = m.group(1).strip()
    return t.replace(" *", "*")


def prototypes(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"CTAPI_STREAM_DEFAULT", "", text)
    out = {}
    for m in re.finditer(r"CTAPI_EXPORT\s+([\w\s\*]+?)\b(ct_\w+)\s*\(([^)]*)\)\s*;", text):
        name, args = m.group(2), m.group(3)
        args = [norm_type(a) for a in args.split(",") if a.strip() and a.strip() != "void"]
        out[name] = {"ret": norm_type(m.group(1)), "args": args}
    return out


def main():
    protos = {}
    for h in ("tacuda.h", "tacuda_device.h", "tacuda_stream.h"):
        protos.update(prototypes(open(os.path.join(REF, "include", h)).read()))
    want = {f"ct_{f}{s}" for f in FAMILIES for s in SUFFIXES} | set(EXTRA)
    want |= {n for n in protos if n.startswith("ct_cdl_")}
    want |= {n for n in protos if "_stream_" in n and not n.startswith("ct_test")}
    fixture = {n: protos[n] for n in sorted(want) if n in protos}
    json.dump(fixture, open(OUT, "w"), indent=0, sort_keys=True)
    print(f"{len(fixture)} prototypes -> {OUT}")


if __name__ == "__main__":
    main()
