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


class BatchPeriodogram:
    """Differentiable periodogram of a batch of S float32 series of N
    samples that share one search (tsamp, widths, period and bins ranges):

        bp = BatchPeriodogram(S, N, tsamp, widths, period_min, ...)
        snrs = bp(x)        # x: (S, N) CUDA tensor -> (S, num_periods, num_widths)

    Forward: every row normalised (rt_grad_batch_normalise), then the batched
    periodogram of the normalised rows (rt_batch_*, the persistent FFA kernel,
    replayed from a hipGraph).  Backward: rt_grad_batch_backward, the adjoint
    of the whole batch in a fixed number of launches.  Per series the results
    equal periodogram_fn's."""

    def __init__(self, nseries, nsamp, tsamp, widths=None, period_min=1.0, period_max=30.0,
                 bins_min=240, bins_max=260, ducy_max=0.20, wtsp=1.5, use_graph=True,
                 group_size=64):
        import ctypes
        from .batch import BatchSearch
        torch = _torch()
        if widths is None:
            widths = generate_width_trials(bins_min, ducy_max=ducy_max, wtsp=wtsp)
        self.widths = np.ascontiguousarray(widths, dtype=np.uint64)
        self.nseries, self.nsamp, self.tsamp = int(nseries), int(nsamp), float(tsamp)
        self.args = (float(period_min), float(period_max), int(bins_min), int(bins_max))
        h = ctypes.c_void_p()
        _lib.call("rt_grad_batch_create", self.nseries, self.nsamp, self.tsamp,
                  _lib.host_ptr(self.widths), self.widths.size, *self.args, ctypes.byref(h))
        self.handle = h
        L = ctypes.c_int64()
        _lib.call("rt_grad_batch_length", h, ctypes.byref(L))
        self.length = L.value
        dev = _lib.device()
        self._y = torch.empty((self.nseries, self.nsamp), dtype=torch.float32, device=dev)
        self._snrs = torch.empty((self.nseries, self.length, self.widths.size),
                                 dtype=torch.float32, device=dev)
        self.batch = BatchSearch([self.nsamp] * self.nseries, self.tsamp, int(bins_min),
                                 int(bins_max), float(period_min), float(period_max),
                                 widths=[self.widths] * self.nseries, deredden=False,
                                 already_normalised=True, group_size=group_size)
        self.batch.bind([self._y[i] for i in range(self.nseries)],
                        [self._snrs[i] for i in range(self.nseries)])
        self.use_graph = use_graph
        self._fn = None

    def grid(self):
        """(periods, foldbins) of the rows of the output."""
        return periodogram_grid(self.nsamp, self.tsamp, *self.args)

    def forward_nograd(self, x):
        """(S/N (S, L, nw) -- a fresh tensor, normalised rows, 1/std per row)."""
        x = x.contiguous()
        if x.shape != (self.nseries, self.nsamp) or x.dtype != _torch().float32:
            raise ValueError(f"x must be a float32 tensor of shape ({self.nseries}, {self.nsamp})")
        inv_std = np.empty(self.nseries, np.float32)
        _lib.call("rt_grad_batch_normalise", self.handle, _lib.ptr(x), _lib.ptr(self._y),
                  _lib.host_ptr(inv_std), _lib.stream())
        self.batch.run(use_graph=self.use_graph)
        return self._snrs.clone(), self._y.clone(), inv_std

    def backward_raw(self, y, inv_std, gsnr):
        torch = _torch()
        gsnr = gsnr.contiguous().float()
        gx = torch.empty_like(y)
        _lib.call("rt_grad_batch_backward", self.handle, _lib.ptr(y), _lib.host_ptr(inv_std),
                  _lib.ptr(gsnr), _lib.ptr(gx), _lib.stream())
        return gx

    def __call__(self, x):
        if self._fn is None:
            torch = _torch()
            bp = self

            class BatchPeriodogramFunction(torch.autograd.Function):
                @staticmethod
                def forward(ctx, xin):
                    snrs, y, inv_std = bp.forward_nograd(xin)
                    ctx.save_for_backward(y)
                    ctx.inv_std = inv_std
                    return snrs

                @staticmethod
                def backward(ctx, gsnr):
                    (y,) = ctx.saved_tensors
                    return bp.backward_raw(y, ctx.inv_std, gsnr)

            self._fn = BatchPeriodogramFunction
        return self._fn.apply(x)

    def close(self):
        if getattr(self, "handle", None) is not None:
            _lib.call("rt_grad_batch_destroy", self.handle)
            self.handle = None
        if getattr(self, "batch", None) is not None:
            self.batch.close()
            self.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
