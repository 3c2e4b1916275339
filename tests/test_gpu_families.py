"""GPU parity of every indicator family against the CPU oracle.

Single series (ct_*_d) at lengths that exercise one segment, several
segments chained by look-back, and partial last segments; dense batches
(ct_*_batch_d) with short rows (one wave per row) and long rows (look-back
inside each row).  Window families are required bit-identical to the
oracle; scan families must be within 1e-5 (tighter than the corpus 1e-4)
and at least 99.9% bit-identical.
"""
import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gpu():
    import tacuda_amd  # noqa: F401
    from tacuda_amd import ops

    return ops


def _orc():
    from oracle import ORC

    return ORC


def _data(n, rows=1, seed=0):
    import tacuda_amd as ta

    return ta.synth_ohlcv(n, rows=rows, seed=seed)


V07 = float(np.float32(0.7))

# name -> (inputs, gpu fn, oracle fn, exact?)
SPECS = {
    "sma20": ("c", lambda o, c: o.sma(c, 20), lambda R, c: R.sma(c, 20), True),
    "sma200": ("c", lambda o, c: o.sma(c, 200), lambda R, c: R.sma(c, 200), True),
    "sum14": ("c", lambda o, c: o.sum(c, 14), lambda R, c: R.sum(c, 14), True),
    "wma20": ("c", lambda o, c: o.wma(c, 20), lambda R, c: R.wma(c, 20), True),
    "var20": ("c", lambda o, c: o.var(c, 20), lambda R, c: R.var(c, 20), True),
    "stddev20": ("c", lambda o, c: o.stddev(c, 20), lambda R, c: R.stddev(c, 20), True),
    "max14": ("c", lambda o, c: o.max(c, 14), lambda R, c: R.max(c, 14), True),
    "min14": ("c", lambda o, c: o.min(c, 14), lambda R, c: R.min(c, 14), True),
    "bbands20": ("c", lambda o, c: o.bbands(c, 20, 2.0, 2.0), lambda R, c: R.bbands(c, 20, 2.0, 2.0), True),
    "bbands20x": ("c", lambda o, c: o.bbands(c, 20, 2.0, 1.5), lambda R, c: R.bbands(c, 20, 2.0, 1.5), True),
    "ema20": ("c", lambda o, c: o.ema(c, 20), lambda R, c: R.ema(c, 20), False),
    "ema200": ("c", lambda o, c: o.ema(c, 200), lambda R, c: R.ema(c, 200), False),
    "ema2": ("c", lambda o, c: o.ema(c, 2), lambda R, c: R.ema(c, 2), False),
    "dema20": ("c", lambda o, c: o.dema(c, 20), lambda R, c: R.dema(c, 20), False),
    "tema20": ("c", lambda o, c: o.tema(c, 20), lambda R, c: R.tema(c, 20), False),
    # the C ABI takes vFactor as float: the oracle gets the same float32 value
    "t3_5": ("c", lambda o, c: o.t3(c, 5, 0.7), lambda R, c: R.t3(c, 5, V07), False),
    "kama30": ("c", lambda o, c: o.kama(c, 30), lambda R, c: R.kama(c, 30), False),
    "kama10_3_20": ("c", lambda o, c: o.kama(c, 10, 3, 20), lambda R, c: R.kama(c, 10, 3, 20), False),
    "rsi14": ("c", lambda o, c: o.rsi(c, 14), lambda R, c: R.rsi(c, 14), False),
    "macd": ("c", lambda o, c: o.macd(c, 12, 26, 9), lambda R, c: R.macd(c, 12, 26, 9), False),
    "macd_sig1": ("c", lambda o, c: o.macd(c, 5, 9, 1), lambda R, c: R.macd(c, 5, 9, 1), False),
    "atr14": ("hlc", lambda o, h, l, c: o.atr(h, l, c, 14), lambda R, h, l, c: R.atr(h, l, c, 14), False),
    "atr1": ("hlc", lambda o, h, l, c: o.atr(h, l, c, 1), lambda R, h, l, c: R.atr(h, l, c, 1), False),
    "natr14": ("hlc", lambda o, h, l, c: o.natr(h, l, c, 14), lambda R, h, l, c: R.natr(h, l, c, 14), False),
    "trange": ("hlc", lambda o, h, l, c: o.trange(h, l, c), lambda R, h, l, c: R.trange(h, l, c), True),
    "stoch14_3": ("hlc", lambda o, h, l, c: o.stochastic(h, l, c, 14, 3),
                  lambda R, h, l, c: R.stoch(h, l, c, 14, 3), True),
    "willr14": ("hlc", lambda o, h, l, c: o.willr(h, l, c, 14), lambda R, h, l, c: R.willr(h, l, c, 14), True),
    "cci14": ("hlc", lambda o, h, l, c: o.cci(h, l, c, 14), lambda R, h, l, c: R.cci(h, l, c, 14), True),
    "plus_dm14": ("hl", lambda o, h, l: o.plus_dm(h, l, 14), lambda R, h, l: R.plus_dm(h, l, 14), False),
    "minus_dm14": ("hl", lambda o, h, l: o.minus_dm(h, l, 14), lambda R, h, l: R.minus_dm(h, l, 14), False),
    "plus_di14": ("hlc", lambda o, h, l, c: o.plus_di(h, l, c, 14), lambda R, h, l, c: R.plus_di(h, l, c, 14), False),
    "minus_di14": ("hlc", lambda o, h, l, c: o.minus_di(h, l, c, 14), lambda R, h, l, c: R.minus_di(h, l, c, 14), False),
    "dx14": ("hlc", lambda o, h, l, c: o.dx(h, l, c, 14), lambda R, h, l, c: R.dx(h, l, c, 14), False),
    "adx14": ("hlc", lambda o, h, l, c: o.adx(h, l, c, 14), lambda R, h, l, c: R.adx(h, l, c, 14), False),
    "aroon14": ("hl", lambda o, h, l: o.aroon(h, l, 14), lambda R, h, l: R.aroon(h, l, 14), True),
    "aroonosc14": ("hl", lambda o, h, l: o.aroonosc(h, l, 14), lambda R, h, l: R.aroonosc(h, l, 14), True),
}


def _run_pair(spec, d, dev):
    kind, gfn, ofn, exact = spec
    ops, R = _gpu(), _orc()
    keys = {"c": ["close"], "hl": ["high", "low"], "hlc": ["high", "low", "close"]}[kind]
    arrs = [d[k] for k in keys]
    tens = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrs]
    got = gfn(ops, *tens)
    torch.cuda.synchronize()
    got = got if isinstance(got, tuple) else (got,)
    got = [g.cpu().numpy() for g in got]
    if arrs[0].ndim == 1:
        want = ofn(R, *arrs)
        want = want if isinstance(want, tuple) else (want,)
    else:
        per = [ofn(R, *[a[r] for a in arrs]) for r in range(arrs[0].shape[0])]
        per = [p if isinstance(p, tuple) else (p,) for p in per]
        want = [np.stack([p[j] for p in per]) for j in range(len(per[0]))]
    return got, want, exact


def _check(name, got, want, exact):
    for j, (g, w) in enumerate(zip(got, want)):
        if exact:
            compare(g, w, tol=0.0, min_exact=1.0, name=f"{name}[{j}]")
        else:
            compare(g, w, tol=1e-5, min_exact=0.999, name=f"{name}[{j}]")


@pytest.mark.parametrize("name", sorted(SPECS))
@pytest.mark.parametrize("n", [1, 37, 256, 1000, 1024, 1025, 70001])
def test_single_series(dev, name, n):
    print(f"single {name} n={n}", flush=True)
    d = _data(n, seed=3)
    got, want, exact = _run_pair(SPECS[name], d, dev)
    _check(f"{name} n={n}", got, want, exact)


@pytest.mark.parametrize("name", sorted(SPECS))
@pytest.mark.parametrize("shape", [(300, 234), (5, 3000)])
def test_dense_batch(dev, name, shape):
    rows, n = shape
    print(f"batch {name} {rows}x{n}", flush=True)
    d = _data(n, rows=rows, seed=5)
    got, want, exact = _run_pair(SPECS[name], d, dev)
    _check(f"{name} {rows}x{n}", got, want, exact)


def _poison(d, prefix=0, holes=(), infs=()):
    out = {k: v.copy() for k, v in d.items()}
    for k in ("open", "high", "low", "close"):
        out[k][:prefix] = np.nan
        for i in holes:
            out[k][i] = np.nan
        for i in infs:
            out[k][i] = np.inf
    return out


@pytest.mark.parametrize("name", sorted(SPECS))
@pytest.mark.parametrize("case", ["prefix7", "prefix1500", "hole", "inf"])
def test_non_finite(dev, name, case):
    print(f"nonfinite {name} {case}", flush=True)
    d = _data(4000, seed=11)
    if case == "prefix7":
        d = _poison(d, prefix=7)
    elif case == "prefix1500":
        d = _poison(d, prefix=1500)
    elif case == "hole":
        d = _poison(d, holes=(100, 2100))
    else:
        d = _poison(d, infs=(300,))
    got, want, exact = _run_pair(SPECS[name], d, dev)
    _check(f"{name} {case}", got, want, exact)


def test_period_longer_than_data(dev):
    ops = _gpu()
    c = torch.arange(10, dtype=torch.float32, device=dev) + 100
    for f in (lambda: ops.ema(c, 20), lambda: ops.sma(c, 11), lambda: ops.rsi(c, 10),
              lambda: ops.macd(c, 12, 26, 9)[0], lambda: ops.t3(c, 5)):
        y = f()
        torch.cuda.synchronize()
        assert torch.isnan(y).all()


def test_invalid_arguments(dev):
    import tacuda_amd as ta

    ops = _gpu()
    c = torch.ones(100, dtype=torch.float32, device=dev)
    for bad in (lambda: ops.ema(c, 0), lambda: ops.sma(c, -1), lambda: ops.t3(c, 1),
                lambda: ops.macd(c, 0, 26, 9)):
        with pytest.raises(ta.TacudaError) as e:
            bad()
        assert e.value.status == 4
    # MAMA is the one TA-Lib moving average MACD does not provide
    with pytest.raises(ta.TacudaError) as e:
        ops.macd(c, 12, 26, 9, ma_type=7)
    assert e.value.status == 6


def test_repeat_launches_reuse_workspace(dev):
    """Epoch-tagged look-back flags: many launches on one workspace agree."""
    ops = _gpu()
    d = _data(200_000, seed=2)
    c = torch.from_numpy(d["close"]).to(dev)
    first = ops.ema(c, 20).cpu().numpy()
    for _ in range(20):
        again = ops.ema(c, 20)
    torch.cuda.synchronize()
    assert np.array_equal(first.view(np.uint32), again.cpu().numpy().view(np.uint32))
