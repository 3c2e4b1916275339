"""GPU parity of the fused long-series scan (ctx_scan_batch_d): EMA, KAMA,
T3 and SMA of each row against the oracle's single-family functions on the
same bars -- within 1e-5 relative and at least 99.9% bit-identical, NaN and
Inf masks identical -- over rows of one segment up to 2^20 bars (1024
segments chained by carries), short and long periods, leading NaNs that
move the seeds across segment boundaries, interior NaN / +Inf / -Inf, rows
shorter than the look-backs and all-NaN rows."""
import numpy as np
import pytest

from tests.parity import compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DEFAULT = dict(ema_period=200, kama_period=30, kama_fast=2, kama_slow=30, t3_period=5,
               t3_vfactor=0.7, sma_period=200)
SHORT = dict(ema_period=2, kama_period=2, kama_fast=3, kama_slow=20, t3_period=2,
             t3_vfactor=0.0, sma_period=1)
LONG = dict(ema_period=1500, kama_period=700, kama_fast=5, kama_slow=50, t3_period=60,
            t3_vfactor=1.2, sma_period=1025)
ONE = dict(DEFAULT, ema_period=1, kama_period=1)
PARAMS = {"default": DEFAULT, "short": SHORT, "long": LONG, "one": ONE}


def _want(R, p, x, name):
    if name == "ema":
        return R.ema(x, p["ema_period"])
    if name == "kama":
        return R.kama(x, p["kama_period"], p["kama_fast"], p["kama_slow"])
    if name == "t3":
        return R.t3(x, p["t3_period"], float(np.float32(p["t3_vfactor"])))
    return R.sma(x, p["sma_period"])


def _check(got, c, p, rows=None):
    from oracle import ORC

    c2 = c.reshape(-1, c.shape[-1])
    rows = list(range(c2.shape[0]) if rows is None else rows)
    for name, t in got.items():
        g = t.reshape(c2.shape)[rows].cpu().numpy()
        for i, r in enumerate(rows):
            compare(g[i], _want(ORC, p, c2[r], name), tol=1e-5, min_exact=0.999,
                    name=f"{name} row {r}")


def _run(c, dev, **p):
    import tacuda_amd as ta

    x = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
    out = ta.ops.scan(x, **p)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("pname", sorted(PARAMS))
@pytest.mark.parametrize("shape", [(5, 700), (3, 5000), (2, 70000), (64, 2100)])
def test_scan_matches_oracle(dev, pname, shape):
    import tacuda_amd as ta

    rows, n = shape
    print(f"scan {pname} {rows}x{n}", flush=True)
    c = ta.synth_ohlcv(n, rows=rows, seed=11)["close"]
    _check(_run(c, dev, **PARAMS[pname]), c, PARAMS[pname])


def test_scan_one_million_bar_row(dev):
    import tacuda_amd as ta

    print("scan 1 x 2^20", flush=True)
    c = ta.synth_ohlcv(1 << 20, seed=0)["close"]
    _check(_run(c, dev, **DEFAULT), c, DEFAULT)


def test_scan_nonfinite_rows(dev):
    import tacuda_amd as ta

    print("scan non-finite", flush=True)
    n = 6000
    c = ta.synth_ohlcv(n, rows=8, seed=5)["close"].copy()
    c[0, :1500] = np.nan            # seeds move across the first segment boundary
    c[1, 3000] = np.nan             # interior NaN
    c[2, 2500] = np.inf
    c[3, 4100] = -np.inf
    c[4, :] = np.nan                # all NaN
    c[5, :1020] = np.nan            # EMA200's seed window straddles two segments
    c[6, 2048:2060] = np.inf        # +Inf inside windows of every family
    c[6, 2100] = -np.inf
    _check(_run(c, dev, **DEFAULT), c, DEFAULT)
    _check(_run(c, dev, **SHORT), c, SHORT)


def test_scan_rows_shorter_than_lookbacks(dev):
    import tacuda_amd as ta

    print("scan short rows", flush=True)
    for n in (1, 5, 25, 150, 199, 200, 201):
        c = ta.synth_ohlcv(n, rows=3, seed=n)["close"]
        _check(_run(c, dev, **DEFAULT), c, DEFAULT)


def test_scan_output_subsets(dev):
    import tacuda_amd as ta

    print("scan subsets", flush=True)
    c = ta.synth_ohlcv(9000, rows=4, seed=8)["close"]
    x = torch.from_numpy(c).to(dev)
    for outs in (("sma",), ("t3",), ("ema", "kama"), ("kama",)):
        got = ta.ops.scan(x, outputs=outs, **DEFAULT)
        torch.cuda.synchronize()
        assert set(got) == set(outs)
        _check(got, c, DEFAULT)


def test_scan_full_shape_sampled_rows(dev):
    """256 x 2^20 bars, the long-scan workload: rows 0, 129 and 255 against
    the oracle; the other rows must equal row r of a smaller launch over the
    same data (rows are independent)."""
    import tacuda_amd as ta

    print("scan 256 x 2^20", flush=True)
    rows, n = 256, 1 << 20
    c = ta.synth_ohlcv(n, rows=rows, seed=0)["close"]
    x = torch.from_numpy(c).to(dev)
    got = ta.ops.scan(x, **DEFAULT)
    torch.cuda.synchronize()
    _check(got, c, DEFAULT, rows=(0, 129, 255))
    part = ta.ops.scan(x[100:104].contiguous(), **DEFAULT)
    torch.cuda.synchronize()
    for k in got:
        assert torch.equal(got[k][100:104].view(torch.int32), part[k].view(torch.int32)), k
    del got, part, x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", [(128, 5000), (16, 40000), (1, 600000)])
def test_family_calls_routed_through_the_scan(dev, shape):
    """ct_ema/kama/t3/sma batch calls on batches the fused scan serves (at
    least 16 rows of two or more 1024-bar segments, 512 segments in all) keep
    the family contract -- the oracle within tolerance, NaN masks equal -- as
    does a single long row, which stays on the look-back kernels."""
    import tacuda_amd as ta

    rows, n = shape
    print(f"scan-routed family calls {rows}x{n}", flush=True)
    c = ta.synth_ohlcv(n, rows=rows, seed=21)["close"].copy()
    c2 = c.reshape(-1, n)
    c2[0, :700] = np.nan
    c2[-1, n // 2] = np.nan
    x = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
    got = {"ema": ta.ops.ema(x, 200), "kama": ta.ops.kama(x, 30), "t3": ta.ops.t3(x, 5, 0.7),
           "sma": ta.ops.sma(x, 200)}
    torch.cuda.synchronize()
    _check(got, c, DEFAULT)
