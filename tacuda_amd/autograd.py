"""Differentiable indicators as torch.autograd.Function over the HIP
forward/adjoint kernels (ctx_*_alpha_fwd_d / _bwd_d, ctx_macd_bwd_d,
ctx_bbands_bwd_d).  Inputs: float32 CUDA tensors of shape [series, bars]
(or [bars]); alpha is a tensor (requires_grad allowed) or a float.

  ema(x, alpha, period)        EMA with weight alpha, TA-Lib SMA seed
  rsi(x, alpha, period)        Wilder RSI with weight alpha
  macd(x, fast, slow, signal)  TA-Lib MACD (line, signal, histogram)
  bbands(x, period, up, dn)    Bollinger bands (upper, middle, lower)

Backward passes save only chunk-boundary carries of the recurrences
(ctx_alpha_ckpt_per_row() per series) and recompute the rest on the GPU.  The adjoint kernels
never read the incoming gradient at warm-up bars (where the outputs are
NaN), so a loss that leaves NaN gradients there needs no masking.
"""
from __future__ import annotations

import torch

from . import call


def _stream():
    return torch.cuda.current_stream().cuda_stream


_CKPT = []


def _ckpt_per_row():
    """Checkpointed states per row and recurrence (one per thread chunk)."""
    if not _CKPT:
        from . import lib

        _CKPT.append(int(lib().ctx_alpha_ckpt_per_row()))
    return _CKPT[0]


def _as2d(x):
    if x.dim() == 1:
        return x.unsqueeze(0), True
    return x, False


def _shape(x):
    return x.shape[1], x.shape[0]  # seriesLen, numSeries


def _alpha_arg(alpha, device):
    """A tensor alpha stays on the device (a float64 copy the kernels read,
    so no host synchronization and the call can be graph-captured); a
    number is passed by value."""
    if torch.is_tensor(alpha):
        if alpha.numel() != 1:
            raise ValueError("alpha must be a scalar tensor")
        return alpha.detach().to(device=device, dtype=torch.float64).reshape(1).contiguous()
    return float(alpha)


class _AlphaRecurrence(torch.autograd.Function):
    fam = ""
    ck_per_lane = 1

    @classmethod
    def forward(cls, ctx, x, alpha, period):
        x2, squeeze = _as2d(x.contiguous())
        n, rows = _shape(x2)
        a = _alpha_arg(alpha, x.device)
        y = torch.empty_like(x2)
        ck = torch.empty(rows * _ckpt_per_row() * cls.ck_per_lane, dtype=torch.float64,
                         device=x.device)
        if torch.is_tensor(a):
            call(f"ctx_{cls.fam}_alpha_fwd_dp", x2.data_ptr(), y.data_ptr(), ck.data_ptr(), n,
                 rows, int(period), a.data_ptr(), _stream())
            ctx.save_for_backward(x2, ck, a)
        else:
            call(f"ctx_{cls.fam}_alpha_fwd_d", x2.data_ptr(), y.data_ptr(), ck.data_ptr(), n,
                 rows, int(period), a, _stream())
            ctx.save_for_backward(x2, ck)
        ctx.meta = (None if torch.is_tensor(a) else a, int(period), squeeze,
                    torch.is_tensor(alpha))
        return y.squeeze(0) if squeeze else y

    @classmethod
    def backward(cls, ctx, gy):
        saved = ctx.saved_tensors
        x2, ck = saved[0], saved[1]
        a, period, squeeze, alpha_is_tensor = ctx.meta
        g2 = (gy.unsqueeze(0) if squeeze else gy).contiguous().float()
        n, rows = _shape(x2)
        gx = torch.empty_like(x2)
        ga = torch.empty(rows, dtype=torch.float64, device=x2.device)
        if len(saved) == 3:
            call(f"ctx_{cls.fam}_alpha_bwd_dp", x2.data_ptr(), g2.data_ptr(), ck.data_ptr(),
                 gx.data_ptr(), ga.data_ptr(), n, rows, period, saved[2].data_ptr(), _stream())
        else:
            call(f"ctx_{cls.fam}_alpha_bwd_d", x2.data_ptr(), g2.data_ptr(), ck.data_ptr(),
                 gx.data_ptr(), ga.data_ptr(), n, rows, period, a, _stream())
        galpha = ga.sum().to(torch.float32) if alpha_is_tensor else None
        return (gx.squeeze(0) if squeeze else gx), galpha, None


class EmaAlpha(_AlphaRecurrence):
    fam = "ema"
    ck_per_lane = 1


class RsiAlpha(_AlphaRecurrence):
    fam = "rsi"
    ck_per_lane = 2


class Macd(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, fast, slow, signal):
        from . import ops

        x2, squeeze = _as2d(x.contiguous())
        m, s, h = ops.macd(x2, fast, slow, signal)
        ctx.meta = (int(fast), int(slow), int(signal), squeeze, x2.shape)
        if squeeze:
            return m.squeeze(0), s.squeeze(0), h.squeeze(0)
        return m, s, h

    @staticmethod
    def backward(ctx, gm, gs, gh):
        fast, slow, signal, squeeze, shape = ctx.meta
        rows, n = shape
        dev = gm.device if gm is not None else (gs if gs is not None else gh).device

        def prep(g):
            if g is None:
                return torch.zeros(shape, dtype=torch.float32, device=dev)
            g = g.unsqueeze(0) if squeeze else g
            return g.contiguous().float()

        gm2, gs2, gh2 = prep(gm), prep(gs), prep(gh)
        gx = torch.empty(shape, dtype=torch.float32, device=dev)
        call("ctx_macd_bwd_d", gm2.data_ptr(), gs2.data_ptr(), gh2.data_ptr(), gx.data_ptr(), n,
             rows, fast, slow, signal, _stream())
        return (gx.squeeze(0) if squeeze else gx), None, None, None


class Bbands(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, period, up, dn):
        from . import ops

        x2, squeeze = _as2d(x.contiguous())
        u, m, l = ops.bbands(x2, period, up, dn)
        ctx.save_for_backward(x2)
        ctx.meta = (int(period), float(up), float(dn), squeeze)
        if squeeze:
            return u.squeeze(0), m.squeeze(0), l.squeeze(0)
        return u, m, l

    @staticmethod
    def backward(ctx, gu, gm, gl):
        (x2,) = ctx.saved_tensors
        period, up, dn, squeeze = ctx.meta
        rows, n = x2.shape

        def prep(g):
            if g is None:
                return torch.zeros_like(x2)
            g = g.unsqueeze(0) if squeeze else g
            return g.contiguous().float()

        gu2, gm2, gl2 = prep(gu), prep(gm), prep(gl)
        gx = torch.empty_like(x2)
        call("ctx_bbands_bwd_d", x2.data_ptr(), gu2.data_ptr(), gm2.data_ptr(), gl2.data_ptr(),
             gx.data_ptr(), n, rows, period, up, dn, _stream())
        return (gx.squeeze(0) if squeeze else gx), None, None, None


def ema(x, alpha, period):
    return EmaAlpha.apply(x, alpha, period)


def rsi(x, alpha, period):
    return RsiAlpha.apply(x, alpha, period)


def macd(x, fast=12, slow=26, signal=9):
    return Macd.apply(x, fast, slow, signal)


def bbands(x, period=20, up=2.0, dn=2.0):
    return Bbands.apply(x, period, up, dn)
