"""Differentiable periodogram: torch.autograd.Function around the HIP forward
and the hand-written adjoint kernels.

    snrs = periodogram_fn(x, tsamp, widths, period_min, period_max, bins_min, bins_max)

x is a float32 CUDA tensor (the dereddened series); the forward normalises it
and computes the periodogram S/N (num_periods x num_widths), exactly as
ffa_search does after dereddening.  The backward recomputes, for every
(step, bins) transform whose rows receive gradient, that transform's forward
and applies the S/N backward (through the first argmax window of every width),
the FFA transpose, the fold and downsampling transposes, then the
normalisation transpose.  Nothing but the normalised series is stored.
"""
import numpy as np

from . import _lib
from .ffautils import generate_width_trials


def _torch():
    import torch
    return torch


def _make_function():
    torch = _torch()

    class PeriodogramFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, tsamp, widths, pmin, pmax, bmin, bmax, normalise):
            if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 1):
                raise ValueError("x must be a 1-D float32 CUDA tensor")
            x = x.contiguous()
            n = x.numel()
            w = np.ascontiguousarray(widths, dtype=np.uint64)
            if normalise:
                y = torch.empty_like(x)
                stats = np.zeros(4, np.float32)
                _lib.call("rt_normalise_ex", _lib.ptr(x), n, _lib.ptr(y), _lib.host_ptr(stats),
                          _lib.stream())
                inv_std = float(1.0 / (np.float64(stats[2]) + np.float64(stats[3])))
            else:
                y = x
                inv_std = 1.0
            from .libcpp import periodogram_device
            _, _, snrs = periodogram_device(y, n, tsamp, w, pmin, pmax, bmin, bmax)
            ctx.save_for_backward(y)
            ctx.params = (tsamp, w, pmin, pmax, bmin, bmax, normalise, inv_std)
            return snrs

        @staticmethod
        def backward(ctx, gsnr):
            (y,) = ctx.saved_tensors
            tsamp, w, pmin, pmax, bmin, bmax, normalise, inv_std = ctx.params
            n = y.numel()
            gsnr = gsnr.contiguous().float()
            gy = torch.empty_like(y)
            _lib.call("rt_periodogram_backward", _lib.ptr(y), n, float(tsamp), _lib.host_ptr(w),
                      w.size, float(pmin), float(pmax), int(bmin), int(bmax), _lib.ptr(gsnr),
                      _lib.ptr(gy), _lib.stream())
            if normalise:
                gx = torch.empty_like(y)
                _lib.call("rt_normalise_backward", _lib.ptr(gy), _lib.ptr(y), n, float(inv_std),
                          _lib.ptr(gx), _lib.stream())
            else:
                gx = gy
            return gx, None, None, None, None, None, None, None

    return PeriodogramFunction


_FN = None


def periodogram_fn(x, tsamp, widths=None, period_min=1.0, period_max=30.0, bins_min=240,
                   bins_max=260, ducy_max=0.20, wtsp=1.5, normalise=True):
    """Differentiable S/N periodogram of a 1-D float32 CUDA tensor."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    if widths is None:
        widths = generate_width_trials(bins_min, ducy_max=ducy_max, wtsp=wtsp)
    return _FN.apply(x, float(tsamp), np.asarray(widths), float(period_min), float(period_max),
                     int(bins_min), int(bins_max), bool(normalise))


def periodogram_grid(nsamp, tsamp, period_min, period_max, bins_min, bins_max):
    """(periods, foldbins) of the rows of periodogram_fn's output."""
    from . import libcpp
    L = libcpp.periodogram_length(nsamp, tsamp, period_min, period_max, bins_min, bins_max)
    P = np.empty(L, np.float64)
    F = np.empty(L, np.uint32)
    _lib.call("rt_periodogram_grid", int(nsamp), float(tsamp), float(period_min),
              float(period_max), int(bins_min), int(bins_max), _lib.host_ptr(P),
              _lib.host_ptr(F))
    return P, F
