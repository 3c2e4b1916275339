"""Benchmark of the indicator core on one MI355X.

Default workload: the device-API 7-indicator pipeline on one GPU-resident
float32 series of 1,000,000 bars (seeded OHLCV, seed 0): SMA(20), EMA(20),
RSI(14), MACD(12,26,9), BBANDS(20,2,2), ATR(14), STOCH(14,3,3).  A step is
one pass of the seven ct_*_d calls, replayed from a hipGraph.  Inputs and
outputs stay in HBM; five buffer sets (300 MB) rotate between steps so the
256 MB Infinity Cache cannot hold the working set.

Prints one JSON line: bars/s for the whole pass, the roofline of the
dominant kernel (algorithmic bytes / its own event-timed duration against
MI355X's 8 TB/s), and the reference's TA-Lib 0.8.3 C timed on the host.

Other workloads (--workload): batch13 (4000 x 234 13-call fused batch).
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


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="pipeline7", choices=["pipeline7"])
    ap.add_argument("--bars", type=int, default=1_000_000)
    ap.add_argument("--sets", type=int, default=5, help="rotated buffer sets")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


# (name, inputs, outputs, call)
def _pipeline(ops):
    return [
        ("sma20", ("c",), 1, lambda h, l, c: ops.sma(c, 20)),
        ("ema20", ("c",), 1, lambda h, l, c: ops.ema(c, 20)),
        ("rsi14", ("c",), 1, lambda h, l, c: ops.rsi(c, 14)),
        ("macd12_26_9", ("c",), 3, lambda h, l, c: ops.macd(c, 12, 26, 9)),
        ("bbands20", ("c",), 3, lambda h, l, c: ops.bbands(c, 20, 2.0, 2.0)),
        ("atr14", ("h", "l", "c"), 1, lambda h, l, c: ops.atr(h, l, c, 14)),
        ("stoch14_3_3", ("h", "l", "c"), 2, lambda h, l, c: ops.stochastic(h, l, c, 14, 3)),
    ]


def _cpu_baseline(d, budget_s):
    """TA-Lib 0.8.3 C (the reference's own host build) on one core."""
    from oracle import REF, ORC, ref

    fam = REF if ref() is not None else ORC
    kind = "reference" if ref() is not None else "port"
    h, l, c = d["high"], d["low"], d["close"]
    calls = [lambda: fam.sma(c, 20), lambda: fam.ema(c, 20), lambda: fam.rsi(c, 14),
             lambda: fam.macd(c, 12, 26, 9), lambda: fam.bbands(c, 20, 2.0, 2.0),
             lambda: fam.atr(h, l, c, 14), lambda: fam.stoch(h, l, c, 14, 3)]
    reps = 0
    t0 = time.perf_counter()
    while True:
        for f in calls:
            f()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 1000:
            break
    bars = reps * len(c)
    return {"value": bars / el, "unit": "bars/s", "cores": 1, "kind": kind,
            "sample": f"{reps} passes of the 7 indicators over the same {len(c)}-bar series, "
                      f"single thread, {el:.1f} s"}


def main():
    a = _args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not "
                                   "supported (series are independent: shard rows by GPU)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import tacuda_amd as ta
    from tacuda_amd import ops

    dev = torch.device("cuda:0")
    n = a.bars
    d = ta.synth_ohlcv(n, seed=0)
    stream = torch.cuda.Stream()
    pipe = _pipeline(ops)
    sets = []
    with torch.cuda.stream(stream):
        for _ in range(a.sets):
            sets.append({k: torch.from_numpy(d[v]).to(dev) for k, v in
                         (("h", "high"), ("l", "low"), ("c", "close"))})
        # eager warm-up on the capture stream reserves its look-back workspace
        for s in sets:
            for _, _, _, f in pipe:
                f(s["h"], s["l"], s["c"])
    stream.synchronize()

    # one graph per buffer set for the whole pipeline, and one per kernel
    graphs, kgraphs, keep = [], {name: [] for name, *_ in pipe}, []
    for s in sets:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            outs = [f(s["h"], s["l"], s["c"]) for _, _, _, f in pipe]
        graphs.append(g)
        keep.append(outs)
        for name, _, _, f in pipe:
            kg = torch.cuda.CUDAGraph()
            with torch.cuda.graph(kg, stream=stream):
                keep.append(f(s["h"], s["l"], s["c"]))
            kgraphs[name].append(kg)
    torch.cuda.synchronize()

    def timed(glist, steps, warm):
        with torch.cuda.stream(stream):
            for i in range(warm):
                glist[i % len(glist)].replay()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            for i in range(steps):
                glist[i % len(glist)].replay()
            e1.record(stream)
        e1.synchronize()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per replay

    ms = timed(graphs, a.steps, a.warmup)
    per_kernel = {}
    for name, ins, nout, _ in pipe:
        kms = timed(kgraphs[name], max(a.steps, 20), a.warmup)
        alg = (len(ins) + nout) * 4 * n
        per_kernel[name] = {"ms": kms, "alg_bytes": alg, "GBs": alg / (kms * 1e-3) / 1e9}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"])
    dk = per_kernel[dom]
    alg_total = (3 + 12) * 4 * n
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(dom)
        except Exception:
            traffic = None
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
        "config": {"workload": "device-API 7-indicator pipeline, 1 x 1,000,000 bars "
                               "(SMA20, EMA20, RSI14, MACD12/26/9, BBANDS20, ATR14, STOCH14/3/3)",
                   "bars": n, "graph": True, "rotated_sets": a.sets,
                   "pass_alg_bytes": alg_total,
                   "pass_GBs": alg_total / (ms * 1e-3) / 1e9,
                   "pass_frac_hbm": alg_total / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                   "kernels": {k: round(v["ms"] * 1e3, 2) for k, v in per_kernel.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["GBs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": dk["GBs"] / PEAK_HBM_GBS, "traffic": traffic},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = _cpu_baseline(d, a.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
