"""FIR whitening filters for a differentiable batched search, and their
optimisation for S/N with Adam.

    y = fir_filter(x, taps)          # x: (S, N), taps: (S, T) float32 CUDA tensors

is differentiable with respect to both x and the taps (hand-written HIP
adjoints, rt_fir_forward / rt_fir_grad_taps).  ``optimise_whitening`` runs the
filter -> normalise -> periodogram chain of every series and updates the taps
by Adam (rt_adam_update) to maximise, per series, the sum over widths of
(the best S/N within +-0.5% of a known period) minus (the median S/N of that
width); the running-median dereddening of ffa_search (time_series.py:95-124)
is the fallback it is judged against.  The reference has no counterpart.
"""
import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def _check(x, taps):
    torch = _torch()
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2):
        raise ValueError("x must be a 2-D (series x samples) float32 CUDA tensor")
    if not (taps.is_cuda and taps.dtype == torch.float32 and taps.dim() == 2):
        raise ValueError("taps must be a 2-D (series x taps) float32 CUDA tensor")
    if taps.shape[0] != x.shape[0]:
        raise ValueError("one row of taps per series is required")


def fir_forward(x, taps, reverse=False):
    """y[s][n] = sum_k taps[s][k] x[s][n + k - (T-1)/2] (no autograd)."""
    torch = _torch()
    _check(x, taps)
    x = x.contiguous()
    taps = taps.contiguous()
    y = torch.empty_like(x)
    _lib.call("rt_fir_forward", _lib.ptr(x), x.shape[0], x.shape[1], _lib.ptr(taps),
              taps.shape[1], 1 if reverse else 0, _lib.ptr(y), _lib.stream())
    return y


def fir_grad_taps(x, g, ntaps):
    """gtaps[s][k] = sum_n g[s][n] x[s][n + k - (T-1)/2]."""
    import ctypes
    torch = _torch()
    x = x.contiguous()
    g = g.contiguous()
    nf = ctypes.c_int64()
    _lib.call("rt_fir_workspace_floats", x.shape[0], x.shape[1], ntaps, ctypes.byref(nf))
    work = torch.empty(max(1, nf.value), dtype=torch.float32, device=x.device)
    gt = torch.empty((x.shape[0], ntaps), dtype=torch.float32, device=x.device)
    _lib.call("rt_fir_grad_taps", _lib.ptr(x), _lib.ptr(g), x.shape[0], x.shape[1], ntaps,
              _lib.ptr(work), _lib.ptr(gt), _lib.stream())
    return gt


def _make_function():
    torch = _torch()

    class FirFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, taps):
            y = fir_forward(x, taps)
            ctx.save_for_backward(x, taps)
            return y

        @staticmethod
        def backward(ctx, gy):
            x, taps = ctx.saved_tensors
            gy = gy.contiguous().float()
            gx = fir_forward(gy, taps, reverse=True) if ctx.needs_input_grad[0] else None
            gt = fir_grad_taps(x, gy, taps.shape[1]) if ctx.needs_input_grad[1] else None
            return gx, gt

    return FirFunction


_FN = None


def fir_filter(x, taps):
    """Differentiable batched FIR filter (see module docstring)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(x, taps)


def identity_taps(nseries, ntaps, device=None):
    torch = _torch()
    if ntaps % 2 == 0:
        raise ValueError("the tap count must be odd")
    t = torch.zeros((nseries, ntaps), dtype=torch.float32, device=device or _lib.device())
    t[:, (ntaps - 1) // 2] = 1.0
    return t


class Adam:
    """Adam on one float32 CUDA parameter tensor (rt_adam_update)."""

    def __init__(self, param, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        torch = _torch()
        self.param = param
        self.lr, self.betas, self.eps = float(lr), betas, float(eps)
        self.m = torch.zeros_like(param)
        self.v = torch.zeros_like(param)
        self.step_count = 0

    def step(self, grad):
        self.step_count += 1
        g = grad.contiguous()
        _lib.call("rt_adam_update", _lib.ptr(self.param), _lib.ptr(g), _lib.ptr(self.m),
                  _lib.ptr(self.v), self.param.numel(), self.lr, float(self.betas[0]),
                  float(self.betas[1]), self.eps, self.step_count, _lib.stream())


def near_period_rows(periods, period, frac=0.005):
    """Indices of the trial periods within +-frac of `period`."""
    p = np.asarray(periods)
    return np.nonzero(np.abs(p - period) <= frac * period)[0]


def whitening_objective(snrs, rows):
    """Sum over widths of (max S/N over `rows`) - (median S/N over all
    periods); snrs is a (periods x widths) tensor."""
    torch = _torch()
    idx = torch.as_tensor(rows, device=snrs.device, dtype=torch.long)
    best = snrs.index_select(0, idx).max(dim=0).values
    med = snrs.median(dim=0).values
    return (best - med).sum()


def _row_index(periods_grid, periods, frac=0.005):
    """(S, R) long tensor of the trial-period rows within +-frac of each
    series' period, padded with each row list's first entry."""
    torch = _torch()
    rows = [near_period_rows(periods_grid, p, frac) for p in periods]
    if any(len(r) == 0 for r in rows):
        raise ValueError("every target period must lie inside the searched period range")
    R = max(len(r) for r in rows)
    idx = np.stack([np.concatenate([r, np.full(R - len(r), r[0])]) for r in rows])
    return torch.as_tensor(idx, dtype=torch.long, device=_lib.device())


def batch_whitening_objective(snrs, row_index):
    """Per series: sum over widths of (max S/N over its rows) - (median S/N
    over all periods); snrs is (S, periods, widths), row_index (S, R)."""
    torch = _torch()
    idx = row_index[:, :, None].expand(-1, -1, snrs.shape[2])
    best = torch.gather(snrs, 1, idx).max(dim=1).values
    med = snrs.median(dim=1).values
    return (best - med).sum(dim=1)


def optimise_whitening(xs, tsamp, periods, steps=50, ntaps=129, lr=2e-3, period_min=0.5,
                       period_max=5.0, bins_min=240, bins_max=260, log=None, batch=None):
    """Optimises one FIR filter per series of xs (S x N float32 CUDA tensor)
    by `steps` Adam steps; periods[s] is the period whose S/N series s should
    maximise.  The whole batch runs through one BatchPeriodogram (batched
    forward and adjoint).  Returns (taps, objective history: steps+1 arrays of
    S values)."""
    from .autograd import BatchPeriodogram
    S, N = xs.shape
    bp = batch or BatchPeriodogram(S, N, tsamp, None, period_min, period_max, bins_min,
                                   bins_max)
    grid, _ = bp.grid()
    rows = _row_index(grid, periods)
    taps = identity_taps(S, ntaps, xs.device)
    opt = Adam(taps, lr=lr)
    history = []
    try:
        for it in range(steps + 1):
            t = taps.detach().requires_grad_(True)
            objs = batch_whitening_objective(bp(fir_filter(xs, t)), rows)
            history.append(objs.detach().cpu().numpy())
            if log:
                log(f"step {it}: mean objective {history[-1].mean():.3f}")
            if it == steps:
                break
            (-objs.sum()).backward()
            opt.step(t.grad)
    finally:
        if batch is None:
            bp.close()
    return taps, history


def detection_significance(snrs, periods_grid, period, exclude=0.02, harmonics=4):
    """Significance of a pulsar at `period` in one periodogram (snrs:
    periods x widths, numpy or tensor): for every width, the best S/N within
    +-0.5% of the period, minus the median, divided by the robust standard
    deviation ((p75 - p25) / 1.349) of the S/N at the trial periods away from
    the period and its harmonics (k*P and P/k for k <= harmonics, +-exclude);
    the largest over widths."""
    S = np.asarray(snrs.detach().cpu().numpy() if hasattr(snrs, "detach") else snrs,
                   dtype=np.float64)
    P = np.asarray(periods_grid)
    near = np.abs(P / period - 1) <= 0.005
    off = np.ones(P.size, bool)
    for k in range(1, harmonics + 1):
        for q in (period * k, period / k):
            off &= np.abs(P / q - 1) > exclude
    best = S[near].max(axis=0)
    p25, med, p75 = np.percentile(S[off], [25, 50, 75], axis=0)
    z = (best - med) / ((p75 - p25) / 1.349)
    return float(z.max())
