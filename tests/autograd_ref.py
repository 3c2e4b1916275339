"""FP64 references of the differentiable indicators, written with plain
torch ops (autograd through the time loop), shared by the finite-difference
check (CPU) and the GPU gradient tests."""
import numpy as np
import torch


def _close(rows, n, seed=0):
    import tacuda_amd as ta

    return ta.synth_ohlcv(n, rows=rows, seed=seed)["close"]


def _weights(shape):
    return np.random.default_rng(2).standard_normal(shape)


def ref_ema(x, alpha, p):
    s = p - 1
    y = [None] * x.shape[1]
    v = x[:, :p].sum(1) / p
    y[s] = v
    for t in range(p, x.shape[1]):
        v = (1 - alpha) * v + alpha * x[:, t]
        y[t] = v
    nan = torch.full_like(x[:, 0], float("nan"))
    return torch.stack([yy if yy is not None else nan for yy in y], 1)


def ref_rsi(x, alpha, p):
    d = x[:, 1:] - x[:, :-1]
    gain, loss = torch.clamp(d, min=0), torch.clamp(-d, min=0)
    G = gain[:, :p].sum(1) / p
    L = loss[:, :p].sum(1) / p
    out = [None] * x.shape[1]
    out[p] = 100 * G / (G + L)
    for t in range(p + 1, x.shape[1]):
        G = (1 - alpha) * G + alpha * gain[:, t - 1]
        L = (1 - alpha) * L + alpha * loss[:, t - 1]
        out[t] = 100 * G / (G + L)
    nan = torch.full_like(x[:, 0], float("nan"))
    return torch.stack([o if o is not None else nan for o in out], 1)


def _coef(p):
    b = (p - 1) / (p + 1)
    k = 1 - b
    if b < 0.5:
        b = 1 - k
    return b, k


def ref_macd(x, fast, slow, sig):
    fb, fk = _coef(fast)
    sb, sk = _coef(slow)
    gb, gk = _coef(sig)
    L1 = slow - 1
    ps = x[:, :slow].sum(1) / slow
    pf = x[:, L1 - fast + 1:L1 + 1].sum(1) / fast
    mv = [None] * x.shape[1]
    mv[L1] = pf - ps
    for t in range(slow, x.shape[1]):
        pf = fb * pf + fk * x[:, t]
        ps = sb * ps + sk * x[:, t]
        mv[t] = pf - ps
    L2 = L1 + sig - 1
    s = sum(mv[L1:L2 + 1]) / sig
    m_out, s_out = [None] * x.shape[1], [None] * x.shape[1]
    m_out[L2], s_out[L2] = mv[L2], s
    for t in range(L2 + 1, x.shape[1]):
        s = mv[t] if sig == 1 else gb * s + gk * mv[t]
        m_out[t], s_out[t] = mv[t], s
    nan = torch.full_like(x[:, 0], float("nan"))
    M = torch.stack([v if v is not None else nan for v in m_out], 1)
    S = torch.stack([v if v is not None else nan for v in s_out], 1)
    return M, S, M - S


def ref_bbands(x, p, up, dn):
    w = x.unfold(1, p, 1)
    mid = w.mean(-1)
    var = ((w - mid.unsqueeze(-1)) ** 2).mean(-1)
    sd = torch.sqrt(var)
    pad = torch.full((x.shape[0], p - 1), float("nan"), dtype=x.dtype)
    U = torch.cat([pad, mid + up * sd], 1)
    M = torch.cat([pad, mid], 1)
    Lo = torch.cat([pad, mid - dn * sd], 1)
    return U, M, Lo


def _loss(outs, ws):
    outs = outs if isinstance(outs, tuple) else (outs,)
    tot = 0.0
    for o, w in zip(outs, ws):
        m = ~torch.isnan(o)
        tot = tot + (torch.where(m, o, torch.zeros_like(o)) * w).sum()
    return tot


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _loss(outs, ws):
    outs = outs if isinstance(outs, tuple) else (outs,)
    tot = 0.0
    for o, w in zip(outs, ws):
        m = ~torch.isnan(o)
        tot = tot + (torch.where(m, o, torch.zeros_like(o)) * w).sum()
    return tot


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = {
    "ema": dict(nout=1, alpha=True),
    "rsi": dict(nout=1, alpha=True),
    "macd": dict(nout=3, alpha=False),
    "bbands": dict(nout=3, alpha=False),
}

ALPHA0 = {"ema": 2.0 / 21.0, "rsi": 1.0 / 14.0, "macd": 0.0, "bbands": 0.0}


def _ref_fn(name):
    return {
        "ema": lambda x, a: ref_ema(x, a, 20),
        "rsi": lambda x, a: ref_rsi(x, a, 14),
        "macd": lambda x, a: ref_macd(x, 12, 26, 9),
        "bbands": lambda x, a: ref_bbands(x, 20, 2.0, 2.0),
    }[name]
