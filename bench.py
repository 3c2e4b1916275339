"""Benchmark of the indicator core on one MI355X.

Headline workload: the 7-indicator pipeline on one GPU-resident float32
series of 1,000,000 bars (seeded OHLCV, seed 0): SMA(20), EMA(20), RSI(14),
MACD(12,26,9), BBANDS(20,2,2), ATR(14), STOCH(14,3,3).  A step is one pass
over the series computing all twelve outputs with the fused launch
(ctx_pack_batch_d), replayed from a hipGraph; the same pass through the
seven drop-in ct_*_d calls is timed beside it.  Inputs and outputs stay in
HBM; five buffer sets rotate between steps so the 256 MB Infinity Cache
cannot hold the working set.

Also measured in the same run (reported under "workloads"):
  batch13  4000 series x 234 bars, the 13 indicators SMA, EMA, WMA, DEMA,
           TEMA, RSI, MACD, BBANDS, ATR, ADX, STOCH, WILLR, CCI (18 outputs)
           in one fused call, hipGraph replay, five rotated buffer sets;
  scan256  256 series x 2^20 bars, EMA(200), KAMA(30), T3(5, 0.7), SMA(200)
           through the fused one-pass scan (ctx_scan_batch_d), with the four
           dense drop-in batch calls timed beside it;
  scan1024 the same scan on 1024 series x 2^20 bars;
  candles  all 61 candlestick patterns, int32 signals, 1024 x 16,384 OHLC;
  ragged   4096 series of lengths U[64, 8192], the 13-indicator pass through
           a tuned, graph-captured plan;
  ticks    one streaming update of 65,536 symbols' EMA/RSI/ATR states;
  host0    ct_sma(20), ct_ema(30), ct_rsi(14) host calls on 1,000,000 bars,
           wall time beside the device-only time (the PCIe share);
  autograd 1024 x 8192, forward+backward vs forward-only time.

Prints one JSON line: bars/s of the headline pass, the roofline of its
dominant kernel (algorithmic bytes per launch / that launch's duration,
timed with HIP events around the kernel on its stream, against MI355X's
8 TB/s), and the reference's TA-Lib 0.8.3 C timed on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
METRIC = "bars/sec and %HBM peak, fused batch + 1M-bar scans; fwd+bwd/fwd time ratio"

PIPE7 = dict(sma_period=20, ema_period=20, rsi_period=14, macd_fast=12, macd_slow=26,
             macd_signal=9, bb_period=20, bb_up=2.0, bb_dn=2.0, atr_period=14, stoch_k=14,
             stoch_d=3)
PIPE7_OUT = ("sma", "ema", "rsi", "macd", "macd_signal", "macd_hist", "bb_upper", "bb_middle",
             "bb_lower", "atr", "stoch_k", "stoch_d")
# algorithmic HBM bytes of each kernel of the two-pass fused pipeline over n
# bars: every input read once, every output written once.  The fold launch
# reads high/low/close and writes the six window outputs (SMA, BBANDS x3,
# STOCH x2); the emit launch reads them again and writes the six recurrence
# outputs (EMA, RSI, MACD x3, ATR); lead, group and rows move a few KB.
PIPE7_KERNEL_BYTES = {
    "k_pipe_fold": lambda n: (3 + 6) * 4 * n,
    "k_pipe_emit": lambda n: (3 + 6) * 4 * n,
    "k_pack_one<8>": lambda n: (3 + 12) * 4 * n,
    "k_pack_one<4>": lambda n: (3 + 12) * 4 * n,
}
# the pass's three parts, each also timed alone as the whole fused call for
# the output subset that needs only it: (inputs read, outputs).  The window
# part alone is one k_pack_win launch; a recurrence part alone runs the two
# passes.
PACK_KERNELS = {
    "windows only": (3, ("sma", "bb_upper", "bb_middle", "bb_lower", "stoch_k", "stoch_d")),
    "close recurrences only": (1, ("ema", "rsi", "macd", "macd_signal", "macd_hist")),
    "ATR only": (3, ("atr",)),
}


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bars", type=int, default=1_000_000)
    ap.add_argument("--sets", type=int, default=5, help="rotated buffer sets")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="",
                    help="comma list of workloads: pipeline7,batch13,scan256,scan1024,candles,ragged,"
                         "ticks,autograd,host0")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MB (counter calibration)")
    return ap.parse_args()


def _drop_in(ops):
    return [
        ("sma20", 1, 1, lambda h, l, c: ops.sma(c, 20)),
        ("ema20", 1, 1, lambda h, l, c: ops.ema(c, 20)),
        ("rsi14", 1, 1, lambda h, l, c: ops.rsi(c, 14)),
        ("macd12_26_9", 1, 3, lambda h, l, c: ops.macd(c, 12, 26, 9)),
        ("bbands20", 1, 3, lambda h, l, c: ops.bbands(c, 20, 2.0, 2.0)),
        ("atr14", 3, 1, lambda h, l, c: ops.atr(h, l, c, 14)),
        ("stoch14_3_3", 3, 2, lambda h, l, c: ops.stochastic(h, l, c, 14, 3)),
    ]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_baseline(d, budget_s):
    """TA-Lib 0.8.3 C (the reference's own host build, oracle/_ref) running
    the 7-indicator pass over the same 1,000,000-bar series: one thread, then
    one series per thread on all the host threads this job may use (the
    hardware threads, capped by OMP_NUM_THREADS).  Inputs are widened and
    outputs allocated before the clock starts (oracle/cpu_bench.cpp)."""
    import ctypes

    from oracle import ORC, ref

    h, l, c = d["high"], d["low"], d["close"]
    n = len(c)
    L = ref()
    if L is None:
        # the port (oracle/oracle.c) through ctypes, single thread
        calls = [lambda: ORC.sma(c, 20), lambda: ORC.ema(c, 20), lambda: ORC.rsi(c, 14),
                 lambda: ORC.macd(c, 12, 26, 9), lambda: ORC.bbands(c, 20, 2.0, 2.0),
                 lambda: ORC.atr(h, l, c, 14), lambda: ORC.stoch(h, l, c, 14, 3)]
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            for f in calls:
                f()
            reps += 1
        el = time.perf_counter() - t0
        return {"value": reps * n / el, "unit": "bars/s", "cores": 1, "kind": "port",
                "cpu_model": _cpu_model(),
                "sample": f"{reps} passes of the 7 indicators over one {n}-bar series, 1 thread"}
    L.ref_bench7.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                     ctypes.POINTER(ctypes.c_double),
                                                     ctypes.POINTER(ctypes.c_longlong)]
    L.ref_bench7.restype = ctypes.c_int

    def run(threads):
        bps, passes = ctypes.c_double(), ctypes.c_longlong()
        rc = L.ref_bench7(h.ctypes.data, l.ctypes.data, c.ctypes.data, n, threads, budget_s,
                          ctypes.byref(bps), ctypes.byref(passes))
        if rc:
            raise RuntimeError(f"ref_bench7 failed: {rc}")
        return bps.value, passes.value

    hw = int(L.ref_hardware_threads())
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(hw, cap) if cap > 0 else hw)
    one, p1 = run(1)
    allc, pa = run(threads)
    cpu = _cpu_model()
    print(f"cpu baseline ({cpu}): 1 thread {one / 1e6:.1f} Mbar/s, {threads} threads "
          f"{allc / 1e6:.1f} Mbar/s", file=sys.stderr, flush=True)
    return {"value": one, "unit": "bars/s", "cores": 1, "kind": "reference", "cpu_model": cpu,
            "sample": f"{p1} passes of the 7 indicators over the seed-0 {n}-bar series on one "
                      f"thread, {budget_s:.0f} s (TA-Lib 0.8.3 C, oracle/cpu_bench.cpp)",
            "all_cores": {"value": allc, "unit": "bars/s", "cores": threads,
                          "hardware_threads": hw,
                          "sample": f"{pa} passes, one series per thread, {budget_s:.0f} s"}}


class Timer:
    """hipGraph capture and event timing on one dedicated stream."""

    def __init__(self, torch):
        self.torch = torch
        self.stream = torch.cuda.Stream()
        self.keep = []

    def graphs(self, sets, fn, reps=1):
        """Eager call per set (reserves workspace), then one graph per set
        holding `reps` back-to-back calls."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for s in sets:
                fn(s)
        self.stream.synchronize()
        out = []
        for s in sets:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self.stream):
                for _ in range(reps):
                    self.keep.append(fn(s))
            out.append(g)
        torch.cuda.synchronize()
        return out

    def time(self, glist, steps, warm):
        """ms per replay: `steps` back-to-back replays bracketed by events on
        the stream the kernels run on.  A second run of `steps` replays with
        an event between replays gives each replay's own duration (median
        and minimum in self.last; those events add a little to each)."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for i in range(warm):
                glist[i % len(glist)].replay()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(self.stream)
            for i in range(steps):
                glist[i % len(glist)].replay()
            e1.record(self.stream)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            ev[0].record(self.stream)
            for i in range(steps):
                glist[i % len(glist)].replay()
                ev[i + 1].record(self.stream)
        ev[-1].synchronize()
        per = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
        mid = len(per) // 2
        self.last = {"mean_ms": e0.elapsed_time(e1) / steps,
                     "median_ms": per[mid] if len(per) % 2 else 0.5 * (per[mid - 1] + per[mid]),
                     "min_ms": per[0], "replays": steps}
        return self.last["mean_ms"]


def _gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def _traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), or
    None when that kernel was not profiled."""
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    return json.load(open(tpath)).get(workload, {}).get(kernel)


def _roofline(workload, kernel, alg_bytes, launch_ms):
    """Roofline of one kernel: algorithmic bytes per launch over its
    event-timed launch duration, against MI355X's 8 TB/s."""
    ach = _gbs(alg_bytes, launch_ms)
    return {"bound": "hbm", "kernel": kernel, "achieved": ach, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": _traffic(workload, kernel),
            "alg_bytes": alg_bytes, "launch_us": round(launch_ms * 1e3, 2)}


def kernel_launch_ms(T, sets, fn, passes):
    """Average event-timed duration (ms) of every kernel the library
    launches in one call: `passes` eager calls on the timer's stream with
    the library's kernel timing on (HIP events recorded on the stream each
    kernel runs on, around that kernel alone)."""
    import tacuda_amd as ta

    torch = T.torch
    with torch.cuda.stream(T.stream):
        for i in range(3):
            fn(sets[i % len(sets)])
        T.stream.synchronize()
        ta.kernel_times()
        ta.kernel_timing(True)
        try:
            for i in range(passes):
                fn(sets[i % len(sets)])
            T.stream.synchronize()
        finally:
            ta.kernel_timing(False)
    acc = {}
    for name, ms in ta.kernel_times():
        acc.setdefault(name, []).append(ms)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def _upload(torch, d, dev, nsets):
    return [{k: torch.from_numpy(d[v]).to(dev) for k, v in
             (("h", "high"), ("l", "low"), ("c", "close"))} for _ in range(nsets)]


def _pipeline7(a, torch, ops, T, dev):
    import tacuda_amd as ta

    n = a.bars
    d = ta.synth_ohlcv(n, seed=0)
    sets = _upload(torch, d, dev, a.sets)
    fused = T.graphs(sets, lambda s: ops.pack(s["h"], s["l"], s["c"], outputs=PIPE7_OUT, **PIPE7))
    ms = T.time(fused, a.steps, a.warmup)
    stats = dict(T.last)
    print(f"pipeline7 fused: {ms * 1e3:.1f} us/pass (median {stats['median_ms'] * 1e3:.1f}, "
          f"min {stats['min_ms'] * 1e3:.1f})", file=sys.stderr, flush=True)
    # the pass's kernels (two passes for a long dense row: k_pipe_fold with
    # the window part, k_pipe_group, k_pipe_emit; one row of 1954 segments
    # needs no k_pipe_rows launch), each timed alone with events around its
    # launch
    launch = kernel_launch_ms(T, sets, lambda s: ops.pack(s["h"], s["l"], s["c"],
                                                          outputs=PIPE7_OUT, **PIPE7), 20)
    print("  kernels: " + ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in launch.items()),
          file=sys.stderr, flush=True)
    kernels = {}
    # ten calls per graph replay: the replay's own overhead is amortized
    for name, (nin, outs) in PACK_KERNELS.items():
        g = T.graphs(sets, lambda s, outs=outs: ops.pack(s["h"], s["l"], s["c"], outputs=outs,
                                                        **PIPE7), reps=10)
        kms = T.time(g, max(a.steps // 5, 10), 2) / 10
        alg = (nin + len(outs)) * 4 * n
        kernels[name] = {"us": kms * 1e3, "alg_bytes": alg, "GBs": _gbs(alg, kms)}
        print(f"  {name}: {kms * 1e3:.1f} us", file=sys.stderr, flush=True)
    drop = {}
    pipe = _drop_in(ops)
    g = T.graphs(sets, lambda s: [f(s["h"], s["l"], s["c"]) for *_, f in pipe])
    drop_ms = T.time(g, a.steps, a.warmup)
    for name, nin, nout, f in pipe:
        gk = T.graphs(sets, lambda s, f=f: f(s["h"], s["l"], s["c"]), reps=10)
        kms = T.time(gk, max(a.steps // 5, 10), 2) / 10
        drop[name] = round(kms * 1e3, 2)
    print(f"pipeline7 drop-in: {drop_ms * 1e3:.1f} us/pass {drop}", file=sys.stderr, flush=True)
    alg_total = (3 + 12) * 4 * n
    return d, {"ms": ms, "stats": stats, "bars": n, "alg_bytes": alg_total, "kernels": kernels,
               "launch_ms": launch, "drop_in_ms": drop_ms, "drop_in_kernels_us": drop}


def _batch13(a, torch, ops, T, dev):
    import tacuda_amd as ta

    rows, n = 4000, 234
    d = ta.synth_ohlcv(n, rows=rows, seed=0)
    sets = _upload(torch, d, dev, a.sets)
    g = T.graphs(sets, lambda s: ops.pack(s["h"], s["l"], s["c"]))
    ms = T.time(g, a.steps, a.warmup)
    stats = dict(T.last)
    bars = rows * n
    alg = (3 + 18) * 4 * bars
    launch = kernel_launch_ms(T, sets, lambda s: ops.pack(s["h"], s["l"], s["c"]), 20)
    dom = max(launch, key=launch.get)
    print(f"batch13 4000x234: {ms * 1e3:.1f} us/pass (median {stats['median_ms'] * 1e3:.1f}, "
          f"min {stats['min_ms'] * 1e3:.1f}); " +
          ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in launch.items()), file=sys.stderr,
          flush=True)
    # one launch computes the whole pass: its algorithmic bytes are the pass's
    return {"ms": ms, "stats": stats, "bars_per_s": bars / (ms * 1e-3), "alg_bytes": alg,
            "GBs": _gbs(alg, ms), "frac_hbm": _gbs(alg, ms) / PEAK_HBM_GBS,
            "shape": [rows, n], "graph": True, "rotated_sets": a.sets,
            "kernels_us": {k: round(v * 1e3, 2) for k, v in launch.items()},
            "roofline": _roofline("batch13", dom, alg, launch[dom])}


def _scan256(a, torch, ops, T, dev):
    """256 x 2^20 close: EMA(200), KAMA(30), T3(5, 0.7), SMA(200).  The fused
    one-pass launch (ctx_scan_batch_d) is the workload's figure; the four
    drop-in family calls are timed beside it."""
    import tacuda_amd as ta

    rows, n = 256, 1 << 20
    dd = ta.synth_ohlcv(n, rows=rows, seed=0)
    c = torch.from_numpy(dd["close"]).to(dev)
    hl = [torch.from_numpy(dd[k]).to(dev) for k in ("high", "low")]
    del dd
    bars = rows * n
    alg = (1 + 4) * 4 * bars
    g = T.graphs([c], lambda x: ops.scan(x))
    fused = T.time(g, max(5, a.steps // 5), 2)
    stats = dict(T.last)
    T.keep.clear()
    torch.cuda.empty_cache()
    print(f"scan256 fused: {fused:.3f} ms/pass", file=sys.stderr, flush=True)
    # per kernel: the one-pass scan reads the close rows once and writes the
    # four outputs
    launch = kernel_launch_ms(T, [c], lambda x: ops.scan(x), 3)
    kbytes = {"k_scan_one": (1 + 4) * 4 * bars}
    kern = {k: {"ms": v, **({"alg_bytes": kbytes[k], "GBs": _gbs(kbytes[k], v),
                             "frac_hbm": _gbs(kbytes[k], v) / PEAK_HBM_GBS}
                            if k in kbytes else {})} for k, v in launch.items()}
    print("  kernels: " + ", ".join(f"{k} {v:.3f} ms" for k, v in launch.items()),
          file=sys.stderr, flush=True)
    torch.cuda.empty_cache()
    calls = {"ema200": lambda x: ops.ema(x, 200), "kama30": lambda x: ops.kama(x, 30),
             "t3_5": lambda x: ops.t3(x, 5, 0.7), "sma200": lambda x: ops.sma(x, 200)}
    per = {}
    total = 0.0
    for name, f in calls.items():
        g = T.graphs([c], f)
        kms = T.time(g, max(5, a.steps // 10), 2)
        per[name] = {"ms": kms, "GBs": _gbs(2 * 4 * rows * n, kms)}
        total += kms
        T.keep.clear()
        torch.cuda.empty_cache()
    print(f"scan256 drop-in 4 calls: {total:.3f} ms/pass", file=sys.stderr, flush=True)
    # the Wilder/MACD family calls on the same rows (dense batches of long
    # rows take the fused pass's two passes)
    fam = {"rsi14": (1, 1, lambda x: ops.rsi(x, 14)),
           "macd12_26_9": (1, 3, lambda x: ops.macd(x, 12, 26, 9)),
           "atr14": (3, 1, lambda x: ops.atr(hl[0], hl[1], x, 14))}
    famr = {}
    for name, (nin, nout, f) in fam.items():
        g = T.graphs([c], f)
        kms = T.time(g, max(5, a.steps // 10), 2)
        famr[name] = {"ms": kms, "GBs": _gbs((nin + nout) * 4 * bars, kms)}
        T.keep.clear()
        torch.cuda.empty_cache()
    print("scan256 family calls: " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in famr.items()),
          file=sys.stderr, flush=True)
    del hl
    return {"ms": fused, "stats": stats, "bars_per_s": bars / (fused * 1e-3), "alg_bytes": alg,
            "GBs": _gbs(alg, fused), "frac_hbm": _gbs(alg, fused) / PEAK_HBM_GBS,
            "shape": [rows, n], "graph": True, "kernels": kern,
            "roofline": _roofline("scan256", "k_scan_one", alg, launch["k_scan_one"]),
            "drop_in_ms": total,
            "drop_in_frac_hbm": _gbs(alg, total) / PEAK_HBM_GBS, "drop_in_kernels": per,
            "family_calls": famr}


def _scan1024(a, torch, ops, T, dev):
    """1024 x 2^20 close through the fused scan (EMA(200), KAMA(30),
    T3(5, 0.7), SMA(200)): the 256 seeded rows of scan256 stacked four
    times (rows are independent, so the cost is that of 1024 distinct
    rows)."""
    import tacuda_amd as ta

    rows, n = 1024, 1 << 20
    base = torch.from_numpy(ta.synth_ohlcv(n, rows=256, seed=0)["close"]).to(dev)
    c = base.repeat(rows // 256, 1).contiguous()
    del base
    bars = rows * n
    alg = (1 + 4) * 4 * bars
    g = T.graphs([c], lambda x: ops.scan(x))
    ms = T.time(g, max(5, a.steps // 10), 1)
    stats = dict(T.last)
    T.keep.clear()
    torch.cuda.empty_cache()
    launch = kernel_launch_ms(T, [c], lambda x: ops.scan(x), 2)
    torch.cuda.empty_cache()
    kbytes = {"k_scan_one": (1 + 4) * 4 * bars}
    kern = {k: {"ms": v, **({"alg_bytes": kbytes[k], "GBs": _gbs(kbytes[k], v),
                             "frac_hbm": _gbs(kbytes[k], v) / PEAK_HBM_GBS}
                            if k in kbytes else {})} for k, v in launch.items()}
    print(f"scan1024 fused: {ms:.3f} ms/pass; kernels: " +
          ", ".join(f"{k} {v:.3f} ms" for k, v in launch.items()), file=sys.stderr, flush=True)
    return {"ms": ms, "stats": stats, "bars_per_s": bars / (ms * 1e-3), "alg_bytes": alg,
            "GBs": _gbs(alg, ms), "frac_hbm": _gbs(alg, ms) / PEAK_HBM_GBS, "shape": [rows, n],
            "graph": True, "data": "the 256 seeded scan256 rows stacked 4x", "kernels": kern,
            "roofline": _roofline("scan1024", "k_scan_one", alg, launch["k_scan_one"])}


def _candles(a, torch, ops, T, dev):
    """All 61 candlestick patterns, int32 signals, 1024 x 16,384 OHLC in one
    launch (ctx_cdl_all_batch_d)."""
    import tacuda_amd as ta

    rows, n = 1024, 16384
    d = ta.synth_ohlcv(n, rows=rows, seed=0)
    x = {k: torch.from_numpy(d[k]).to(dev) for k in ("open", "high", "low", "close")}
    g = T.graphs([x], lambda s: ops.cdl_all(s["open"], s["high"], s["low"], s["close"]))
    ms = T.time(g, max(5, a.steps // 5), 2)
    T.keep.clear()
    torch.cuda.empty_cache()
    bars = rows * n
    alg = (4 + 61) * 4 * bars
    launch = kernel_launch_ms(T, [x], lambda s: ops.cdl_all(s["open"], s["high"], s["low"],
                                                            s["close"]), 5)
    torch.cuda.empty_cache()
    kname, kms = max(launch.items(), key=lambda kv: kv[1])
    print(f"candles 61 x 1024x16384: {ms:.3f} ms/pass, {kname} {kms:.3f} ms", file=sys.stderr,
          flush=True)
    return {"ms": ms, "bars_per_s": bars / (ms * 1e-3), "alg_bytes": alg, "GBs": _gbs(alg, ms),
            "frac_hbm": _gbs(alg, ms) / PEAK_HBM_GBS, "shape": [rows, n], "graph": True,
            "patterns": 61, "dtype": "int32",
            "kernel": {"name": kname, "ms": kms, "GBs": _gbs(alg, kms),
                       "frac_hbm": _gbs(alg, kms) / PEAK_HBM_GBS}}


def _ragged(a, torch, ops, T, dev):
    """4096 series, lengths uniform in [64, 8192] (seed 1), the 13-indicator
    fused pass through a tuned, graph-captured plan."""
    import numpy as np

    import tacuda_amd as ta

    lengths = np.random.default_rng(1).integers(64, 8193, size=4096)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    d = ta.synth_ohlcv(int(lengths.max()), rows=len(lengths), seed=0)
    cat = {k: np.concatenate([d[k][r, :m] for r, m in enumerate(lengths)]) for k in
           ("high", "low", "close")}
    del d
    Tn = [torch.from_numpy(cat[k]).to(dev) for k in ("high", "low", "close")]
    lay = ops.Ragged(off)
    with torch.cuda.stream(T.stream):
        plan = ops.Plan(*Tn, layout=lay, tune=True)
        info = plan.info()
        for _ in range(3):
            plan.launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        T.stream.synchronize()
        reps = max(5, a.steps // 5)
        e0.record(T.stream)
        for _ in range(reps):
            plan.launch()
        e1.record(T.stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    bars = int(lengths.sum())
    alg = (3 + 18) * 4 * bars
    print(f"ragged 4096 x U[64,8192]: {ms * 1e3:.1f} us/pass {info}", file=sys.stderr, flush=True)
    plan.close()
    lay.close()
    return {"ms": ms, "bars": bars, "bars_per_s": bars / (ms * 1e-3), "alg_bytes": alg,
            "GBs": _gbs(alg, ms), "frac_hbm": _gbs(alg, ms) / PEAK_HBM_GBS, "plan": info}


def _host_config0(a, torch, ops, dev):
    """The reference's plain host-API case: ct_sma(20), ct_ema(30), ct_rsi(14)
    on one host float32 series of 1,000,000 bars (seed 0), each call taking
    host arrays (pooled device buffers, pinned staging).  Reported per call:
    the wall time of the host call, the device-only time of the same
    computation (ct_*_d on resident tensors, events), and their difference,
    the host<->device copy and synchronization share."""
    import numpy as np

    import tacuda_amd as ta

    L = ta.lib()
    n = 1_000_000
    c = np.ascontiguousarray(ta.synth_ohlcv(n, seed=0)["close"])
    out = np.empty_like(c)
    x = torch.from_numpy(c).to(dev)
    y = torch.empty_like(x)
    s = torch.cuda.current_stream()
    res = {}
    for name, p in (("sma", 20), ("ema", 30), ("rsi", 14)):
        host = getattr(L, f"ct_{name}")
        devf = getattr(L, f"ct_{name}_d")
        for _ in range(3):
            assert host(c.ctypes.data, out.ctypes.data, n, p, None) == 0
        reps = 30
        wall = []
        for _ in range(reps):
            t0 = time.perf_counter()
            assert host(c.ctypes.data, out.ctypes.data, n, p, None) == 0
            wall.append(time.perf_counter() - t0)
        for _ in range(3):
            devf(x.data_ptr(), y.data_ptr(), n, p, s.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            devf(x.data_ptr(), y.data_ptr(), n, p, s.cuda_stream)
        e1.record(s)
        e1.synchronize()
        dms = e0.elapsed_time(e1) / reps
        hms = float(np.median(wall)) * 1e3
        res[f"{name}{p}"] = {"host_call_ms_median": hms, "host_call_ms_min": min(wall) * 1e3,
                             "device_ms": dms, "copy_and_sync_ms": hms - dms,
                             "host_bars_per_s": n / (hms * 1e-3)}
    print("host config0 1M bars: " + ", ".join(
        f"{k} host {v['host_call_ms_median']:.3f} ms (device {v['device_ms']:.3f})"
        for k, v in res.items()), file=sys.stderr, flush=True)
    return res


def _ticks(a, torch, dev):
    """Streaming tick: 65,536 symbols' EMA(20), RSI(14) and ATR(14) states
    advanced one bar per tick (three batch handles, GPU execution)."""
    import numpy as np

    import tacuda_amd as ta
    from tacuda_amd.stream import StreamBatch

    S, hist, ticks = 65536, 64, 200
    d = ta.synth_ohlcv(hist + ticks, rows=S, seed=0)
    bs = [StreamBatch("ema", S, 20, exec="gpu"), StreamBatch("rsi", S, 14, exec="gpu"),
          StreamBatch("atr", S, 14, exec="gpu")]

    def step(b, sl):
        kw = dict(high=np.ascontiguousarray(d["high"][:, sl]),
                  low=np.ascontiguousarray(d["low"][:, sl]),
                  close=np.ascontiguousarray(d["close"][:, sl]))
        return b.append(**({"close": kw["close"]} if b.family != "atr" else kw))

    for b in bs:
        step(b, slice(0, hist))
    lat = []
    for t in range(hist, hist + ticks):
        t0 = time.perf_counter()
        for b in bs:
            step(b, slice(t, t + 1))
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e6
    # the same ticks from device memory (states stay in HBM, one sync per tick)
    dd = {k: torch.from_numpy(np.ascontiguousarray(d[k][:, hist:])).to(dev)
          for k in ("high", "low", "close")}
    cols = [{k: dd[k][:, i].contiguous() for k in dd} for i in range(ticks)]
    outs = [torch.empty(S, device=dev) for _ in range(3)]
    dlat = []
    for b in bs:
        b.reset()
    for i in range(ticks):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bs[0].append_d(close=cols[i]["close"], outs=[outs[0]])
        bs[1].append_d(close=cols[i]["close"], outs=[outs[1]])
        bs[2].append_d(high=cols[i]["high"], low=cols[i]["low"], close=cols[i]["close"],
                       outs=[outs[2]])
        torch.cuda.synchronize()
        dlat.append(time.perf_counter() - t0)
    dlat = np.array(dlat[5:]) * 1e6
    for b in bs:
        b.close()
    res = {"symbols": S, "indicators": ["EMA20", "RSI14", "ATR14"], "ticks": ticks,
           "host_api_us_p50": float(np.percentile(lat, 50)),
           "host_api_us_p99": float(np.percentile(lat, 99)),
           "host_api_ticks_per_s": float(1e6 / np.median(lat)),
           "device_api_us_p50": float(np.percentile(dlat, 50)),
           "device_api_us_p99": float(np.percentile(dlat, 99)),
           "device_api_ticks_per_s": float(1e6 / np.median(dlat))}
    print(f"ticks 65536 x (EMA,RSI,ATR): host p50 {res['host_api_us_p50']:.0f} us, "
          f"device p50 {res['device_api_us_p50']:.0f} us", file=sys.stderr, flush=True)
    return res


def _autograd_ratio(a, torch, dev):
    """1024 x 8192 close: EMA(alpha), MACD(12,26,9), RSI(alpha, 14),
    BBANDS(20, 2); loss = sum(out * w), w ~ N(0,1) seed 2.  Forward-only
    time vs forward+backward time."""
    import tacuda_amd as ta
    from tacuda_amd import autograd as ag

    rows, n = 1024, 8192
    c = torch.from_numpy(ta.synth_ohlcv(n, rows=rows, seed=0)["close"]).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    w = torch.randn(8, rows, n, device=dev, generator=g)
    ae = torch.tensor(2.0 / 21.0, device=dev, requires_grad=True)
    ar = torch.tensor(1.0 / 14.0, device=dev, requires_grad=True)

    def loss_of(x):
        outs = [ag.ema(x, ae, 20), *ag.macd(x, 12, 26, 9), ag.rsi(x, ar, 14),
                *ag.bbands(x, 20, 2.0, 2.0)]
        return sum((torch.nan_to_num(o, nan=0.0) * w[i]).sum() for i, o in enumerate(outs))

    def run(backward):
        x = c.detach().requires_grad_(backward)
        if backward:
            loss_of(x).backward()
        else:
            with torch.no_grad():
                loss_of(x)

    res = {}
    for name, bw in (("fwd", False), ("fwd_bwd", True)):
        for _ in range(3):
            run(bw)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            run(bw)
        e1.record()
        e1.synchronize()
        res[name + "_ms"] = e0.elapsed_time(e1) / reps
    res["ratio"] = res["fwd_bwd_ms"] / res["fwd_ms"]
    res["shape"] = [rows, n]
    print(f"autograd 1024x8192: fwd {res['fwd_ms']:.3f} ms, fwd+bwd {res['fwd_bwd_ms']:.3f} ms, "
          f"ratio {res['ratio']:.2f}", file=sys.stderr, flush=True)
    return res


def main():
    a = _args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not "
                                   "supported (series are independent: shard rows by GPU)"}))
        sys.exit(2)
    import torch

    from tacuda_amd import ops

    dev = torch.device("cuda:0")
    only = set(filter(None, a.only.split(","))) or {"pipeline7", "batch13", "scan256", "scan1024",
                                                    "candles", "ragged", "ticks", "autograd",
                                                    "host0"}
    T = Timer(torch)
    work = {}
    d = p7 = None
    if "pipeline7" in only:
        d, p7 = _pipeline7(a, torch, ops, T, dev)
    if "batch13" in only:
        work["batch13_4000x234"] = _batch13(a, torch, ops, T, dev)
    T.keep.clear()
    torch.cuda.empty_cache()
    if "scan256" in only:
        work["scan256_ema200_kama30_t3_sma200"] = _scan256(a, torch, ops, T, dev)
    if "scan1024" in only:
        work["scan1024_ema200_kama30_t3_sma200"] = _scan1024(a, torch, ops, T, dev)
        torch.cuda.empty_cache()
    if "candles" in only:
        work["candles_61_1024x16384"] = _candles(a, torch, ops, T, dev)
    if "ragged" in only:
        work["ragged_4096_U64_8192"] = _ragged(a, torch, ops, T, dev)
        torch.cuda.empty_cache()
    if "host0" in only:
        work["host_api_1M_sma20_ema30_rsi14"] = _host_config0(a, torch, ops, dev)
    if "ticks" in only:
        work["stream_ticks_65536"] = _ticks(a, torch, dev)
    if "autograd" in only:
        work["autograd_1024x8192"] = _autograd_ratio(a, torch, dev)
        torch.cuda.empty_cache()
    if a.copy_mb:
        x = torch.empty(a.copy_mb << 18, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
    if p7 is None:
        print(json.dumps({"workloads": work}))
        return
    n, ms = p7["bars"], p7["ms"]
    launch = p7["launch_ms"]
    dom = max(launch, key=launch.get)
    line = {
        "metric": METRIC,
        "value": n / (ms * 1e-3),
        "unit": "bars/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 io, f64 recurrences",
        "data": "synthetic seeded OHLCV (geometric random walk, seed 0)",
        "config": {"workload": "7-indicator pipeline, 1 x 1,000,000 bars (SMA20, EMA20, RSI14, "
                               "MACD12/26/9, BBANDS20, ATR14, STOCH14/3/3), fused launch",
                   "bars": n, "graph": True, "rotated_sets": a.sets,
                   "step_ms_median": p7["stats"]["median_ms"],
                   "step_ms_min": p7["stats"]["min_ms"],
                   "pass_alg_bytes": p7["alg_bytes"],
                   "pass_GBs": _gbs(p7["alg_bytes"], ms),
                   "pass_frac_hbm": _gbs(p7["alg_bytes"], ms) / PEAK_HBM_GBS,
                   "kernels_us": {k: round(v * 1e3, 2) for k, v in launch.items()},
                   "kernels_sum_us": round(sum(launch.values()) * 1e3, 2),
                   "parts_alone_us": {k: round(v["us"], 2) for k, v in p7["kernels"].items()},
                   "drop_in_7call_ms": p7["drop_in_ms"],
                   "drop_in_bars_per_s": n / (p7["drop_in_ms"] * 1e-3),
                   "drop_in_kernels_us": p7["drop_in_kernels_us"]},
        "roofline": _roofline("pipeline7", dom, PIPE7_KERNEL_BYTES.get(dom, lambda n: 0)(n),
                              launch[dom]),
        "workloads": work,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = _cpu_baseline(d, a.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
