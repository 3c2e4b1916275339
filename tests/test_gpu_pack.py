"""GPU parity of the fused multi-indicator launch (ctx_pack_batch_d).

Every output of the pack is checked against the oracle's single-family
function on the same bars: window outputs (SMA, WMA, BBANDS, STOCH, WILLR,
CCI) bit-identical, recurrence outputs (EMA, DEMA, TEMA, RSI, MACD, ATR,
ADX) within 1e-5 and at least 99.9% bit-identical, NaN masks identical.
"""
import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EXACT = {"sma", "wma", "bb_upper", "bb_middle", "bb_lower", "stoch_k", "stoch_d", "willr", "cci"}

DEFAULT = dict(sma_period=20, ema_period=20, wma_period=20, dema_period=20, tema_period=20,
               rsi_period=14, macd_fast=12, macd_slow=26, macd_signal=9, bb_period=20, bb_up=2.0,
               bb_dn=2.0, atr_period=14, adx_period=14, stoch_k=14, stoch_d=3, willr_period=14,
               cci_period=20)
# short and degenerate periods: period-1 contracts, swapped MACD periods,
# signal 1, asymmetric bands, ADX 2, STOCH 1/1
EDGE = dict(sma_period=1, ema_period=2, wma_period=1, dema_period=2, tema_period=2, rsi_period=1,
            macd_fast=26, macd_slow=12, macd_signal=1, bb_period=1, bb_up=2.0, bb_dn=1.5,
            atr_period=1, adx_period=2, stoch_k=1, stoch_d=1, willr_period=1, cci_period=1)
LONG = dict(sma_period=200, ema_period=50, wma_period=100, dema_period=50, tema_period=50,
            rsi_period=30, macd_fast=19, macd_slow=39, macd_signal=9, bb_period=50, bb_up=2.5,
            bb_dn=2.5, atr_period=50, adx_period=30, stoch_k=50, stoch_d=9, willr_period=60,
            cci_period=40)
PERIOD1 = dict(DEFAULT, ema_period=1, dema_period=1, tema_period=1)
# window look-backs that need 192- and 1024-bar halos in the window kernel
MID = dict(DEFAULT, sma_period=150, wma_period=33, bb_period=120, stoch_k=100, stoch_d=5,
           willr_period=90, cci_period=64)
HUGE = dict(DEFAULT, sma_period=1000, wma_period=700, bb_period=600, stoch_k=900, stoch_d=9,
            willr_period=1025, cci_period=513)
PARAMS = {"default": DEFAULT, "edge": EDGE, "long": LONG, "ema1": PERIOD1, "mid": MID,
          "huge": HUGE}


def _want(R, p, h, l, c, name):
    if name == "sma":
        return R.sma(c, p["sma_period"])
    if name == "ema":
        return R.ema(c, p["ema_period"])
    if name == "wma":
        return R.wma(c, p["wma_period"])
    if name == "dema":
        return R.dema(c, p["dema_period"])
    if name == "tema":
        return R.tema(c, p["tema_period"])
    if name == "rsi":
        return R.rsi(c, p["rsi_period"])
    if name.startswith("macd"):
        m = R.macd(c, p["macd_fast"], p["macd_slow"], p["macd_signal"])
        return m[{"macd": 0, "macd_signal": 1, "macd_hist": 2}[name]]
    if name.startswith("bb_"):
        b = R.bbands(c, p["bb_period"], float(np.float32(p["bb_up"])), float(np.float32(p["bb_dn"])))
        return b[{"bb_upper": 0, "bb_middle": 1, "bb_lower": 2}[name]]
    if name == "atr":
        return R.atr(h, l, c, p["atr_period"])
    if name == "adx":
        return R.adx(h, l, c, p["adx_period"])
    if name.startswith("stoch"):
        return R.stoch(h, l, c, p["stoch_k"], p["stoch_d"])[0 if name == "stoch_k" else 1]
    if name == "willr":
        return R.willr(h, l, c, p["willr_period"])
    if name == "cci":
        return R.cci(h, l, c, p["cci_period"])
    raise KeyError(name)


def _run(dev, d, params, outputs=None, close_only=False):
    import tacuda_amd as ta
    from oracle import ORC

    outputs = outputs or ta.PACK_OUTPUTS
    T = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")}
    if close_only:
        got = ta.ops.pack(None, None, T["close"], outputs=outputs, **params)
    else:
        got = ta.ops.pack(T["high"], T["low"], T["close"], outputs=outputs, **params)
    torch.cuda.synchronize()
    h, l, c = d["high"], d["low"], d["close"]
    for name in outputs:
        g = got[name].cpu().numpy()
        if c.ndim == 1:
            w = _want(ORC, params, h, l, c, name)
        else:
            w = np.stack([_want(ORC, params, h[r], l[r], c[r], name) for r in range(c.shape[0])])
        if name in EXACT:
            compare(g, w, tol=0.0, min_exact=1.0, name=name)
        else:
            compare(g, w, tol=1e-5, min_exact=0.999, name=name)


def _data(n, rows=1, seed=0):
    import tacuda_amd as ta

    return ta.synth_ohlcv(n, rows=rows, seed=seed)


@pytest.mark.parametrize("pname", sorted(PARAMS))
@pytest.mark.parametrize("n", [1, 37, 256, 1000, 1025, 70001])
def test_pack_single_series(dev, pname, n):
    print(f"pack single {pname} n={n}", flush=True)
    _run(dev, _data(n, seed=11), PARAMS[pname])


@pytest.mark.parametrize("pname", ["default", "edge", "long"])
@pytest.mark.parametrize("shape", [(4000, 234), (300, 257), (5, 3000), (2, 300000)])
def test_pack_dense_batch(dev, pname, shape):
    rows, n = shape
    print(f"pack batch {pname} {rows}x{n}", flush=True)
    _run(dev, _data(n, rows=rows, seed=13), PARAMS[pname])


@pytest.mark.parametrize("subset", [
    ("ema",), ("dema",), ("tema", "rsi"), ("macd_hist",), ("atr",), ("adx",),
    ("sma", "adx"), ("stoch_d", "willr"), ("cci", "tema", "atr"),
])
@pytest.mark.parametrize("n", [300, 20000])
def test_pack_subsets(dev, subset, n):
    print(f"pack subset {subset} n={n}", flush=True)
    _run(dev, _data(n, seed=17), DEFAULT, outputs=subset)


def test_pack_close_only_inputs(dev):
    print("pack close-only", flush=True)
    outs = ("sma", "ema", "wma", "dema", "tema", "rsi", "macd", "macd_signal", "macd_hist",
            "bb_upper", "bb_middle", "bb_lower")
    _run(dev, _data(5000, seed=19), DEFAULT, outputs=outs, close_only=True)


def _poison(d, prefix=0, holes=(), infs=()):
    out = {k: v.copy() for k, v in d.items()}
    for k in ("open", "high", "low", "close"):
        out[k][..., :prefix] = np.nan
        for i in holes:
            out[k][..., i] = np.nan
        for i in infs:
            out[k][..., i] = np.inf
    return out


@pytest.mark.parametrize("case", ["prefix7", "prefix1500", "hole", "inf"])
def test_pack_non_finite(dev, case):
    print(f"pack non-finite {case}", flush=True)
    d = _data(6000, seed=23)
    kw = {"prefix7": dict(prefix=7), "prefix1500": dict(prefix=1500),
          "hole": dict(holes=(3000,)), "inf": dict(infs=(2500,))}[case]
    _run(dev, _poison(d, **kw), DEFAULT)


def test_pack_period_longer_than_series(dev):
    print("pack period > n", flush=True)
    _run(dev, _data(40, seed=29), LONG)


def test_pack_repeat_launches_agree(dev):
    """Back-to-back launches on one stream reuse the look-back workspace."""
    import tacuda_amd as ta

    print("pack repeat", flush=True)
    d = _data(200000, seed=31)
    T = [torch.from_numpy(d[k]).to(dev) for k in ("high", "low", "close")]
    first = ta.ops.pack(*T)
    for _ in range(3):
        again = ta.ops.pack(*T)
    torch.cuda.synchronize()
    for k in first:
        a, b = first[k].cpu().numpy(), again[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k


# Long dense rows whose recurrences are EMA, RSI, MACD and ATR (periods of 2
# or more) with window look-backs within 64 bars run as two passes: per
# segment composites, a per-row carry, then the replay.
PIPE_OUT = ("sma", "ema", "rsi", "macd", "macd_signal", "macd_hist", "bb_upper", "bb_middle",
            "bb_lower", "atr", "stoch_k", "stoch_d", "willr", "cci", "wma")
PIPE_REC = ("ema", "rsi", "macd", "macd_signal", "macd_hist", "atr")


def _pipe_launches(rec):
    """Fused-pass calls in a kernel-timer record: the first kernel of each,
    k_pipe_solo (one launch, few rows) or k_pipe_fold (the two passes)."""
    return sum(1 for k, _ in rec if k in ("k_pipe_solo", "k_pipe_fold"))


@pytest.mark.parametrize("shape", [(1, 8192), (1, 8193), (1, 70001), (3, 20037), (40, 9000)])
def test_pack_two_pass_rows(dev, shape):
    rows, n = shape
    print(f"pack two-pass {rows}x{n}", flush=True)
    _run(dev, _data(n, rows=rows, seed=37), DEFAULT, outputs=PIPE_OUT)


def test_pack_two_pass_row_carry_launch(dev):
    """Many segments (8 rows of 2^20 + 3 bars, 33 groups of 64 segments per
    row): the state entering each group comes from the k_pipe_rows launch
    instead of the emit waves' own scans."""
    import tacuda_amd as ta

    print("pack two-pass 8x1048579 (row carry launch)", flush=True)
    d = _data((1 << 20) + 3, rows=8, seed=39)
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        _run(dev, d, DEFAULT, outputs=PIPE_REC)
    finally:
        ta.kernel_timing(False)
    names = [k for k, _ in ta.kernel_times()]
    assert names == ["k_pipe_fold", "k_pipe_group", "k_pipe_rows", "k_pipe_emit"]


def _bits_equal(got, want, what):
    for k in want:
        a, b = got[k].cpu().numpy(), want[k].cpu().numpy()
        assert np.array_equal(np.isnan(a), np.isnan(b)), (what, k)
        ok = ~np.isnan(a)
        assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), (what, k)


@pytest.mark.parametrize("outs", [("rsi",), ("macd", "macd_signal", "macd_hist"), ("atr",),
                                  ("ema",)])
def test_pack_one_launch_epochs_across_shapes_and_graphs(dev, outs):
    """Passes of one recurrence group run as one launch whose look-back
    records carry per-row epochs that each launch's last tiles advance.
    Calls of four shapes on one stream, a graph of one of them replayed
    between eager calls of the others, and repeats of every shape give the
    bits of each shape's first call, which match the oracle."""
    import tacuda_amd as ta

    print(f"pack one-launch epochs across shapes {outs}", flush=True)
    PIPE_OUT = outs
    shapes = [(1, 70001), (3, 20037), (5, 9000), (1, 300000)]
    data = {sh: _data(sh[1], rows=sh[0], seed=71 + i) for i, sh in enumerate(shapes)}
    ins = {sh: tuple(torch.from_numpy(np.ascontiguousarray(data[sh][k])).to(dev)
                     for k in ("high", "low", "close")) for sh in shapes}
    s = torch.cuda.Stream()
    first = {}
    with torch.cuda.stream(s):
        for sh in shapes:
            out = ta.ops.pack(*ins[sh], outputs=PIPE_OUT, **DEFAULT)
            first[sh] = {k: v.clone() for k, v in out.items()}
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gout = ta.ops.pack(*ins[shapes[1]], outputs=PIPE_OUT, **DEFAULT)
    for rep in range(3):
        with torch.cuda.stream(s):
            g.replay()
            got = {k: v.clone() for k, v in gout.items()}
            others = [(sh, ta.ops.pack(*ins[sh], outputs=PIPE_OUT, **DEFAULT))
                      for sh in (shapes[0], shapes[3], shapes[2])]
        s.synchronize()
        _bits_equal(got, first[shapes[1]], f"replay {rep}")
        for sh, out in others:
            _bits_equal(out, first[sh], f"eager {sh} after replay {rep}")
    for sh in shapes:
        _run(dev, data[sh], DEFAULT, outputs=PIPE_OUT)


@pytest.mark.parametrize("pname", ["long", "ema2"])
def test_pack_two_pass_periods(dev, pname):
    params = {"long": LONG, "ema2": dict(DEFAULT, ema_period=2, rsi_period=2, atr_period=2,
                                         macd_fast=2, macd_slow=3, macd_signal=2)}[pname]
    print(f"pack two-pass periods {pname}", flush=True)
    _run(dev, _data(30000, rows=2, seed=41), params, outputs=PIPE_REC)


@pytest.mark.parametrize("subset", [("ema",), ("rsi",), ("macd_signal",), ("atr",),
                                    ("ema", "atr", "sma"), ("rsi", "macd_hist", "cci"),
                                    ("sma", "bb_upper", "bb_lower", "stoch_k", "ema"),
                                    ("bb_middle", "cci", "rsi"), ("bb_lower", "wma", "atr")])
def test_pack_two_pass_subsets(dev, subset):
    print(f"pack two-pass subset {subset}", flush=True)
    _run(dev, _data(17000, rows=2, seed=43), DEFAULT, outputs=subset)


@pytest.mark.parametrize("case", ["prefix1500", "prefix9000", "hole", "inf", "edge"])
def test_pack_two_pass_non_finite(dev, case):
    print(f"pack two-pass non-finite {case}", flush=True)
    d = _data(20000, seed=47)
    kw = {"prefix1500": dict(prefix=1500), "prefix9000": dict(prefix=9000),
          "hole": dict(holes=(10000,)), "inf": dict(infs=(12000,)),
          "edge": dict(holes=(511, 512, 4095))}[case]
    _run(dev, _poison(d, **kw), DEFAULT, outputs=PIPE_OUT)


@pytest.mark.parametrize("case", ["plain", "prefix", "hole", "period1"])
@pytest.mark.parametrize("shape", [(16, 16500), (20, 30011), (2, (1 << 20) + 5)])
def test_family_batches_routed_through_two_passes(dev, case, shape):
    """ct_rsi_batch_d / ct_macd_batch_d / ct_atr_batch_d on dense batches of
    at least 16 long rows and 512 segments, or on fewer rows holding 4096
    segments, take the fused pass's two passes (period 1 keeps the look-back
    kernels): the oracle's results either way."""
    import tacuda_amd as ta
    from oracle import ORC

    rows, n = shape
    print(f"routed families {case} {rows}x{n}", flush=True)
    d = _data(n, rows=rows, seed=59)
    if case == "prefix":
        d = _poison(d, prefix=600)
    elif case == "hole":
        d = _poison(d, holes=(n // 2,))
    p = 1 if case == "period1" else 14
    h, l, c = (torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close"))
    torch.cuda.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        rsi = ta.ops.rsi(c, p)
        atr = ta.ops.atr(h, l, c, p)
        m = ta.ops.macd(c, 12, 26, 9)
    finally:
        ta.kernel_timing(False)
    torch.cuda.synchronize()
    folds = _pipe_launches(ta.kernel_times())
    assert folds == (1 if case == "period1" else 3), folds  # MACD's periods are >= 2
    H, L, C = d["high"], d["low"], d["close"]
    for r in range(rows):
        compare(rsi[r].cpu().numpy(), ORC.rsi(C[r], p), tol=1e-5, min_exact=0.999,
                name=f"rsi row {r}")
        compare(atr[r].cpu().numpy(), ORC.atr(H[r], L[r], C[r], p), tol=1e-5, min_exact=0.999,
                name=f"atr row {r}")
        w = ORC.macd(C[r], 12, 26, 9)
        for i in range(3):
            compare(m[i][r].cpu().numpy(), w[i], tol=1e-5, min_exact=0.999, name=f"macd{i} row {r}")


def test_single_long_row_macd_routed(dev):
    """One row of 600,000 bars (1172 segments): MACD, RSI and ATR each take
    the fused pass's one-launch form; the oracle's results."""
    import tacuda_amd as ta
    from oracle import ORC

    print("routed families 1x600000", flush=True)
    d = _data(600000, seed=61)
    h, l, c = (torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close"))
    torch.cuda.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        m = ta.ops.macd(c, 12, 26, 9)
        rsi = ta.ops.rsi(c, 14)
        atr = ta.ops.atr(h, l, c, 14)
    finally:
        ta.kernel_timing(False)
    torch.cuda.synchronize()
    assert [k for k, _ in ta.kernel_times()] == ["k_pipe_solo"] * 3
    w = ORC.macd(d["close"], 12, 26, 9)
    for i in range(3):
        compare(m[i].cpu().numpy(), w[i], tol=1e-5, min_exact=0.999, name=f"macd{i}")
    compare(rsi.cpu().numpy(), ORC.rsi(d["close"], 14), tol=1e-5, min_exact=0.999, name="rsi")
    compare(atr.cpu().numpy(), ORC.atr(d["high"], d["low"], d["close"], 14), tol=1e-5,
            min_exact=0.999, name="atr")


def test_single_long_row_ema_rsi_routed(dev):
    """One row of 2^20 bars (2048 segments, 512 tiles: the one-launch form's
    longest row but one tile): EMA, RSI and ATR each take the one-launch
    form; the oracle's results."""
    import tacuda_amd as ta
    from oracle import ORC

    print("routed families 1x1048576", flush=True)
    d = _data(1 << 20, seed=67)
    h, l, c = (torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close"))
    torch.cuda.synchronize()
    ta.kernel_times()
    ta.kernel_timing(True)
    try:
        ema = ta.ops.ema(c, 20)
        rsi = ta.ops.rsi(c, 14)
        atr = ta.ops.atr(h, l, c, 14)
    finally:
        ta.kernel_timing(False)
    torch.cuda.synchronize()
    assert [k for k, _ in ta.kernel_times()] == ["k_pipe_solo"] * 3
    compare(ema.cpu().numpy(), ORC.ema(d["close"], 20), tol=1e-5, min_exact=0.999, name="ema")
    compare(rsi.cpu().numpy(), ORC.rsi(d["close"], 14), tol=1e-5, min_exact=0.999, name="rsi")
    compare(atr.cpu().numpy(), ORC.atr(d["high"], d["low"], d["close"], 14), tol=1e-5,
            min_exact=0.999, name="atr")


def test_pack_two_pass_captured_before_any_eager_call(dev):
    """Captured with periods that never ran eagerly, the pass's steady-wave
    multiplier table comes from k_pipe_qt inside the graph; an eager call
    afterwards uses the host-computed table.  Replays and the eager call
    agree bit for bit, and both agree with the oracle."""
    import tacuda_amd as ta

    print("pack two-pass capture-first 1x300000", flush=True)
    params = dict(DEFAULT, ema_period=23, rsi_period=17, macd_fast=11, macd_slow=29,
                  macd_signal=7, atr_period=19)
    d = _data(300000, rows=1, seed=43)
    h, l, c = (torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close"))
    s = torch.cuda.Stream()
    assert ta.lib().ctx_workspace_reserve(s.cuda_stream, 8 << 20) == 0
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        got = ta.ops.pack(h, l, c, outputs=PIPE_REC, **params)
    reps = []
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        reps.append({k: v.clone() for k, v in got.items()})
    want = ta.ops.pack(h, l, c, outputs=PIPE_REC, **params)
    torch.cuda.synchronize()
    for r in reps:
        for k in PIPE_REC:
            a, b = r[k].cpu().numpy(), want[k].cpu().numpy()
            assert np.array_equal(np.isnan(a), np.isnan(b)), k
            ok = ~np.isnan(a)
            assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32)), k
    _run(dev, d, params, outputs=PIPE_REC)
