"""The FP64 reference gradients of the differentiable indicators against
central finite differences in FP64 (h = 1e-6), on the CPU."""
import numpy as np
import pytest
import torch

from tests.autograd_ref import ALPHA0, CASES, _close, _loss, _ref_fn, _weights


@pytest.mark.parametrize("name", ["ema", "rsi", "macd", "bbands"])
def test_fp64_reference_against_finite_differences(name):
    """The reference gradient itself: central differences, h = 1e-6."""
    rows, n = 2, 400
    c = _close(rows, n, seed=3).astype(np.float64)
    nout = CASES[name]["nout"]
    ws = [torch.tensor(_weights((rows, n))) for _ in range(nout)]
    a0 = ALPHA0[name]
    x = torch.tensor(c, requires_grad=True)
    a = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
    _loss(_ref_fn(name)(x, a), ws).backward()
    h = 1e-6
    rng = np.random.default_rng(4)
    for _ in range(6):
        r, t = int(rng.integers(rows)), int(rng.integers(n))
        xp, xm = c.copy(), c.copy()
        xp[r, t] += h
        xm[r, t] -= h
        fd = (_loss(_ref_fn(name)(torch.tensor(xp), a0), ws) -
              _loss(_ref_fn(name)(torch.tensor(xm), a0), ws)).item() / (2 * h)
        g = x.grad[r, t].item()
        assert abs(fd - g) <= 1e-4 * max(1.0, abs(g)), (name, r, t, fd, g)
    if CASES[name]["alpha"]:
        fd = (_loss(_ref_fn(name)(x.detach(), a0 + h), ws) -
              _loss(_ref_fn(name)(x.detach(), a0 - h), ws)).item() / (2 * h)
        assert abs(fd - a.grad.item()) <= 1e-5 * abs(a.grad.item()), (fd, a.grad.item())
