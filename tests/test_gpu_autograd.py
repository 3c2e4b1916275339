"""Differentiable indicators: gradients of loss = sum(out * w), w ~ N(0,1)
(seed 2), w.r.t. the prices and the smoothing weight alpha, against an
FP64 reference written with plain torch ops on the CPU (autograd through the
time loop), which is itself checked by central finite differences in FP64
(h = 1e-6).  Bound: relative error <= 1e-5 (norm-wise for the price
gradient)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests.autograd_ref import (ALPHA0, CASES, _close, _loss, _ref_fn,  # noqa: E402
                                 _rel, _weights)

ROWS, N = 16, 8192
TOL = 1e-5


def _gpu_fn(name):
    from tacuda_amd import autograd as ag

    return {
        "ema": lambda x, a: ag.ema(x, a, 20),
        "rsi": lambda x, a: ag.rsi(x, a, 14),
        "macd": lambda x, a: ag.macd(x, 12, 26, 9),
        "bbands": lambda x, a: ag.bbands(x, 20, 2.0, 2.0),
    }[name]


ALPHA0 = {"ema": 2.0 / 21.0, "rsi": 1.0 / 14.0, "macd": 0.0, "bbands": 0.0}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradients_against_fp64_reference(dev, name):
    _check_against_reference(dev, name, ROWS, N)


# Rows of 40 and 500 bars leave most of a row's 512 threads without bars; 777
# and 4099 bars end inside a thread's chunk and inside a wave; 32768 is the
# longest row (64 bars a thread, the largest LDS columns).
@pytest.mark.parametrize("name,n", [(name, n) for name in sorted(CASES)
                                    for n in (40, 500, 777, 4099)] +
                         [("ema", 16384), ("rsi", 16384), ("ema", 32768), ("rsi", 32768),
                          ("macd", 32768), ("bbands", 32768)])
def test_gradients_row_lengths(dev, name, n):
    _check_against_reference(dev, name, 3, n, seed=n)


def _check_against_reference(dev, name, rows, n, seed=0):
    print(f"autograd {name} rows {rows} bars {n}", flush=True)
    spec = CASES[name]
    c = _close(rows, n, seed=seed)
    ROWS, N = rows, n
    nout = spec["nout"]
    # the weights and alpha the GPU receives are float32: the FP64 reference
    # takes the same rounded values, so the bound measures the kernels'
    # arithmetic (at 500 bars the alpha gradient's cancellation turns the
    # weights' float32 rounding alone into a 1.2e-5 relative change)
    ws = [_weights((ROWS, N)).astype(np.float32).astype(np.float64) for _ in range(nout)]
    a32 = float(np.float32(ALPHA0[name]))
    # FP64 reference gradient
    xr = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    ar = torch.tensor(a32, dtype=torch.float64, requires_grad=True)
    ref_out = _ref_fn(name)(xr, ar)
    Lr = _loss(ref_out, [torch.tensor(w) for w in ws])
    Lr.backward()
    gx_ref = xr.grad.numpy()
    ga_ref = ar.grad.item() if (spec["alpha"] and ar.grad is not None) else None
    # GPU
    xg = torch.tensor(c, device=dev, requires_grad=True)
    ag_ = torch.tensor(ALPHA0[name], dtype=torch.float32, device=dev, requires_grad=True)
    out = _gpu_fn(name)(xg, ag_ if spec["alpha"] else None)
    outs = out if isinstance(out, tuple) else (out,)
    ref_outs = ref_out if isinstance(ref_out, tuple) else (ref_out,)
    for o, r in zip(outs, ref_outs):
        o = o.detach().cpu().numpy().astype(np.float64)
        r = r.detach().numpy()
        assert np.array_equal(np.isnan(o), np.isnan(r))
        ok = ~np.isnan(r)
        rel = np.abs(o[ok] - r[ok]) / np.maximum(1.0, np.abs(r[ok]))
        assert rel.max() <= 1e-5, f"{name} forward rel {rel.max()}"
    Lg = _loss(outs, [torch.tensor(w, dtype=torch.float32, device=dev) for w in ws])
    Lg.backward()
    gx = xg.grad.cpu().numpy()
    err = _rel(gx, gx_ref)
    assert err <= TOL, f"{name}: price gradient rel error {err:.3g}"
    if ga_ref is not None:
        ga = ag_.grad.item()
        assert abs(ga - ga_ref) / abs(ga_ref) <= TOL, f"{name}: alpha gradient {ga} vs {ga_ref}"


def test_full_size_forward_backward_runs(dev):
    """1024 x 8192, the differentiable workload's shape: finite gradients."""
    from tacuda_amd import autograd as ag

    print("autograd full size", flush=True)
    c = torch.tensor(_close(1024, 8192, seed=1), device=dev, requires_grad=True)
    a = torch.tensor(2.0 / 21.0, device=dev, requires_grad=True)
    outs = [ag.ema(c, a, 20)] + list(ag.macd(c)) + [ag.rsi(c, torch.tensor(1 / 14, device=dev), 14)]
    outs += list(ag.bbands(c, 20, 2.0, 2.0))
    loss = sum(torch.nan_to_num(o, nan=0.0).sum() for o in outs)
    loss.backward()
    assert torch.isfinite(c.grad).all() and torch.isfinite(a.grad)


@pytest.mark.parametrize("name", sorted(CASES))
def test_warmup_gradients_are_not_read(dev, name):
    """An incoming gradient that is NaN at the warm-up bars (where the
    outputs are NaN) gives the same price and alpha gradients as one that is
    zero there."""
    print(f"autograd warm-up {name}", flush=True)
    spec = CASES[name]
    c = _close(4, 1000, seed=3)
    res = []
    for fill in (0.0, float("nan")):
        xg = torch.tensor(c, device=dev, requires_grad=True)
        ag_ = torch.tensor(ALPHA0[name], dtype=torch.float32, device=dev, requires_grad=True)
        out = _gpu_fn(name)(xg, ag_ if spec["alpha"] else None)
        outs = out if isinstance(out, tuple) else (out,)
        grads = []
        for i, o in enumerate(outs):
            w = torch.tensor(np.random.default_rng(10 + i).standard_normal(tuple(o.shape)),
                             dtype=torch.float32, device=dev)
            grads.append(torch.where(torch.isnan(o), torch.full_like(w, fill), w))
        torch.autograd.backward(outs, grads)
        res.append((xg.grad.cpu().numpy(), ag_.grad.item() if spec["alpha"] else 0.0))
    assert np.isfinite(res[1][0]).all()
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


@pytest.mark.parametrize("name", ["ema", "rsi"])
def test_device_alpha_matches_host_alpha_and_captures(dev, name):
    """A tensor alpha is read on the device (ctx_*_alpha_*_dp): the same
    outputs and gradients as the same alpha passed as a number, with no host
    synchronization, so the forward can be captured into a graph; updating
    the alpha tensor in place between replays changes the replayed outputs."""
    from tacuda_amd import autograd as ag

    print(f"autograd device alpha {name}", flush=True)
    c = torch.tensor(_close(8, 3000, seed=4), device=dev)
    a0 = float(np.float32(ALPHA0[name]))
    fn = _gpu_fn(name)
    xa = c.clone().requires_grad_(True)
    at = torch.tensor(a0, dtype=torch.float32, device=dev, requires_grad=True)
    ya = fn(xa, at)
    xb = c.clone().requires_grad_(True)
    yb = fn(xb, a0)
    w = torch.tensor(np.random.default_rng(5).standard_normal(tuple(c.shape)),
                     dtype=torch.float32, device=dev)
    (torch.nan_to_num(ya, nan=0.0) * w).sum().backward()
    (torch.nan_to_num(yb, nan=0.0) * w).sum().backward()
    assert torch.equal(torch.isnan(ya), torch.isnan(yb))
    assert torch.equal(torch.nan_to_num(ya), torch.nan_to_num(yb))
    assert torch.equal(xa.grad, xb.grad)
    # capture the forward with a tensor alpha; replay after changing alpha
    alpha = torch.tensor(a0, dtype=torch.float32, device=dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        fn(c, alpha)  # eager first (allocations outside the capture)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        yg = fn(c, alpha)
    for a in (a0, float(np.float32(a0 * 0.5)), a0):
        alpha.fill_(a)
        g.replay()
        torch.cuda.synchronize()
        want = fn(c, a)
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(yg), torch.nan_to_num(want)), a
