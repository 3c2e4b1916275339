"""Persistent tuned plans (ctx_pack_plan_create): the autotuner's choice is
one of the tile configurations, graph replays equal each other bit for bit
and the oracle within tolerance, with eager work and input rewrites between
replays (the graph reads the live buffers)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _check_all(got, d, params):
    from oracle import ORC
    from tests.parity import compare
    from tests.test_gpu_pack import EXACT, _want

    h, l, c = d["high"], d["low"], d["close"]
    for name, t in got.items():
        g = t.cpu().numpy()
        if c.ndim == 1:
            w = _want(ORC, params, h, l, c, name)
        else:
            w = np.stack([_want(ORC, params, h[r], l[r], c[r], name) for r in range(c.shape[0])])
        if name in EXACT:
            compare(g, w, tol=0.0, min_exact=1.0, name=name)
        else:
            compare(g, w, tol=1e-5, min_exact=0.999, name=name)


@pytest.mark.parametrize("shape", [(4000, 234), (1, 200000), (8, 30000)])
def test_plan_replays_match_oracle(dev, shape):
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT

    rows, n = shape
    print(f"plan {rows}x{n}", flush=True)
    d = ta.synth_ohlcv(n, rows=rows, seed=41)
    T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")]
    plan = ta.ops.Plan(*T, tune=True, **DEFAULT)
    info = plan.info()
    assert (info["bars_per_lane"], info["waves_per_group"]) in {
        (4, 4), (4, 2), (4, 1), (4, 8), (8, 2), (8, 1), (8, 4)}
    assert info["tuned_us"] > 0 and info["tuned_us"] <= info["default_us"] * 1.0001
    first = {k: v.clone() for k, v in plan.launch().items()}
    torch.cuda.synchronize()
    _check_all(first, d, DEFAULT)
    for _ in range(3):
        junk = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
        out = plan.launch()
        torch.cuda.synchronize()
        for k in first:
            # bit patterns: NaN warm-ups compare equal
            assert torch.equal(out[k].view(torch.int32), first[k].view(torch.int32)), k
    # new data in the same input buffers: the next replay reads it
    d2 = ta.synth_ohlcv(n, rows=rows, seed=43)
    for t, k in zip(T, ("high", "low", "close")):
        t.copy_(torch.from_numpy(np.ascontiguousarray(d2[k])).to(dev))
    out = plan.launch()
    torch.cuda.synchronize()
    _check_all(out, d2, DEFAULT)
    plan.close()
    del junk


def test_plan_over_ragged_layout(dev):
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT
    from tests.test_gpu_ragged import _layout, _rows
    from oracle import ORC
    from tests.parity import compare
    from tests.test_gpu_pack import EXACT, _want

    print("plan ragged", flush=True)
    lengths = list(np.random.default_rng(3).integers(64, 3000, size=300))
    off, cat = _layout(lengths, seed=5)
    lay = ta.ops.Ragged(off)
    T = [torch.from_numpy(cat[k]).to(dev) for k in ("high", "low", "close")]
    plan = ta.ops.Plan(*T, layout=lay, tune=True, **DEFAULT)
    out = plan.launch()
    torch.cuda.synchronize()
    from tests.test_gpu_ragged import _cat_want

    for name, t in out.items():
        want = _cat_want(off, lambda a, b, d: _want(ORC, DEFAULT, a, b, d, name),
                         cat["high"], cat["low"], cat["close"])
        if name in EXACT:
            compare(t.cpu().numpy(), want, tol=0.0, min_exact=1.0, name=name)
        else:
            compare(t.cpu().numpy(), want, tol=1e-5, min_exact=0.999, name=name)
    plan.close()
    lay.close()


@pytest.mark.parametrize("shape", [(300, 234), (1, 100000), (3, 40000)])
def test_user_graph_capture_of_the_fused_pass(dev, shape):
    """ctx_pack_batch_d captured by the caller on a stream whose look-back
    workspace an eager call reserved: every replay equals the eager pass bit
    for bit, with eager work and cleared outputs between replays."""
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT

    rows, n = shape
    print(f"user graph {rows}x{n}", flush=True)
    d = ta.synth_ohlcv(n, rows=rows, seed=7)
    T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")]
    T = [t if rows > 1 else t.reshape(-1) for t in T]
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        eager = {k: v.clone() for k, v in ta.ops.pack(*T, **DEFAULT).items()}
    st.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        cap = ta.ops.pack(*T, **DEFAULT)
    torch.cuda.synchronize()
    for _ in range(3):
        junk = torch.randn(1 << 20, device=dev).sum()
        for v in cap.values():
            v.fill_(0.0)
        g.replay()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(cap[k].view(torch.int32), eager[k].view(torch.int32)), k
    del junk, g


@pytest.mark.parametrize("shape", [(1, 200000), (6, 20000)])
def test_two_pass_plan_and_user_capture(dev, shape):
    """The two-pass fused pass (long dense rows, EMA/RSI/MACD/ATR with short
    windows) as a tuned plan and under the caller's own graph capture: every
    replay equals the oracle, and the eager call bit for bit."""
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT, PIPE_OUT

    rows, n = shape
    print(f"two-pass plan/capture {rows}x{n}", flush=True)
    d = ta.synth_ohlcv(n, rows=rows, seed=53)
    T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")]
    T = [t if rows > 1 else t.reshape(-1) for t in T]
    plan = ta.ops.Plan(*T, outputs=PIPE_OUT, tune=True, **DEFAULT)
    first = {k: v.clone() for k, v in plan.launch().items()}
    torch.cuda.synchronize()
    _check_all(first, d, DEFAULT)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        eager = {k: v.clone() for k, v in ta.ops.pack(*T, outputs=PIPE_OUT, **DEFAULT).items()}
    st.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        cap = ta.ops.pack(*T, outputs=PIPE_OUT, **DEFAULT)
    torch.cuda.synchronize()
    for _ in range(3):
        junk = torch.randn(1 << 20, device=dev).sum()
        for v in cap.values():
            v.fill_(0.0)
        g.replay()
        out = plan.launch()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(cap[k].view(torch.int32), eager[k].view(torch.int32)), k
            assert torch.equal(out[k].view(torch.int32), first[k].view(torch.int32)), k
    plan.close()
    del junk, g


def test_graph_replay_after_workspace_growth(dev):
    """A graph captured on a stream keeps the look-back workspace it was
    captured with: a larger eager call on the same stream afterwards moves
    the stream to a bigger workspace (the old one stays allocated), and the
    graph's replays still equal the eager pass bit for bit; the large call
    matches the oracle on sampled rows."""
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT

    print("workspace growth after capture", flush=True)
    L = ta.lib()
    d = ta.synth_ohlcv(234, rows=300, seed=71)
    T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")]
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        eager = {k: v.clone() for k, v in ta.ops.pack(*T, **DEFAULT).items()}
    st.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        cap = ta.ops.pack(*T, **DEFAULT)
    torch.cuda.synchronize()
    before = L.ct_device_buffer_pool_allocation_count()
    big = ta.synth_ohlcv(2340, rows=4000, seed=72)
    B = [torch.from_numpy(np.ascontiguousarray(big[k])).to(dev) for k in ("high", "low", "close")]
    with torch.cuda.stream(st):
        out = ta.ops.pack(*B, **DEFAULT)
    st.synchronize()
    assert L.ct_device_buffer_pool_allocation_count() > before, "workspace did not grow"
    for _ in range(3):
        for v in cap.values():
            v.fill_(0.0)
        g.replay()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(cap[k].view(torch.int32), eager[k].view(torch.int32)), k
    sub = {k: v[[0, 1999, 3999]] for k, v in big.items()}
    _check_all({k: v[[0, 1999, 3999]] for k, v in out.items()}, sub, DEFAULT)
    del g


def test_plan_launches_on_two_streams(dev):
    """Replays of one plan on two streams, issued back to back without host
    synchronization, are ordered by the plan (they share its look-back
    workspace): every replay's outputs equal the first replay's and the
    oracle."""
    import tacuda_amd as ta
    from tests.test_gpu_pack import DEFAULT

    print("plan on two streams", flush=True)
    d = ta.synth_ohlcv(234, rows=4000, seed=73)
    T = [torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in ("high", "low", "close")]
    plan = ta.ops.Plan(*T, tune=False, **DEFAULT)
    first = {k: v.clone() for k, v in plan.launch().items()}
    torch.cuda.synchronize()
    sub = {k: v[[0, 17, 3999]] for k, v in d.items()}
    _check_all({k: v[[0, 17, 3999]] for k, v in first.items()}, sub, DEFAULT)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    snaps = []
    for i in range(6):
        s = sa if i % 2 == 0 else sb
        with torch.cuda.stream(s):
            out = plan.launch()
            snaps.append({k: v.clone() for k, v in out.items()})
    torch.cuda.synchronize()
    for snap in snaps:
        for k in first:
            assert torch.equal(snap[k].view(torch.int32), first[k].view(torch.int32)), k
    plan.close()


def test_tuned_ragged_plan_4096_series(dev):
    """The fused uneven-length batch at its real shape: 4096 series of
    lengths uniform in [64, 8192] (seed 1), the 13 indicators through a
    tuned, graph-captured plan; 32 sampled series (the shortest and longest
    among them) equal the oracle on all 18 outputs."""
    import tacuda_amd as ta
    from oracle import ORC
    from tests.parity import compare
    from tests.test_gpu_pack import DEFAULT, EXACT, _want

    print("tuned ragged plan 4096 x U[64,8192]", flush=True)
    lengths = np.random.default_rng(1).integers(64, 8193, size=4096)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    d = ta.synth_ohlcv(int(lengths.max()), rows=len(lengths), seed=0)
    cat = {k: np.concatenate([d[k][r, :m] for r, m in enumerate(lengths)]) for k in
           ("high", "low", "close")}
    T = [torch.from_numpy(cat[k]).to(dev) for k in ("high", "low", "close")]
    lay = ta.ops.Ragged(off)
    plan = ta.ops.Plan(*T, layout=lay, tune=True, **DEFAULT)
    out = {k: v.cpu().numpy() for k, v in plan.launch().items()}
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    sample = set(rng.choice(4096, size=30, replace=False).tolist())
    sample |= {int(np.argmin(lengths)), int(np.argmax(lengths))}
    for r in sorted(sample):
        sl = slice(off[r], off[r + 1])
        h, l, c = (cat[k][sl] for k in ("high", "low", "close"))
        for name, g in out.items():
            w = _want(ORC, DEFAULT, h, l, c, name)
            if name in EXACT:
                compare(g[sl], w, tol=0.0, min_exact=1.0, name=f"{name} series {r}")
            else:
                compare(g[sl], w, tol=1e-5, min_exact=0.999, name=f"{name} series {r}")
    plan.close()
    lay.close()
