"""Folding of a time series at a candidate period (folding.py:6-85 of the
reference): same arguments, checks, messages and output shapes; the
downsampling, scaling, row sums and vertical downsampling run on the GPU."""
import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def downsample_vertical(X, factor):
    """Downsample every column of a 2-D array by a real factor
    (folding.py:6-18)."""
    X = np.asarray(X)
    m, bins = X.shape
    if not factor > 1:
        raise ValueError("factor must be > 1")
    if not factor < m:
        raise ValueError("factor must be strictly smaller than the number of input lines")
    n = int(np.floor(m / factor))
    d = _lib.to_device(np.ascontiguousarray(X, np.float32))
    out = _lib.empty((max(n, 1), bins))
    _lib.call("rt_downsample_vertical", _lib.ptr(d), m, bins, float(factor), _lib.ptr(out),
              _lib.stream())
    return _lib.to_numpy(out)[:n]


def fold(ts, period, bins, subints=None):
    """Fold a TimeSeries at `period` seconds into `bins` phase bins
    (folding.py:21-85): (subints, bins) array, or (bins,) when subints == 1
    or when only one period fits."""
    if period > ts.length:
        raise ValueError("Period exceeds data length")
    tbin = period / bins
    if not tbin > ts.tsamp:
        raise ValueError("Bin width is shorter than sampling time")
    if subints is not None:
        subints = int(subints)
        if not subints >= 1:
            raise ValueError("subints must be >= 1 or None")
        full_periods = ts.length / period
        if subints > full_periods:
            raise ValueError(
                f"subints ({subints}) exceeds the number of signal periods that fit in the "
                f"data ({full_periods})")
    factor = tbin / ts.tsamp
    x = np.ascontiguousarray(ts.data, np.float32)
    if not ((factor > 1.0) and (factor <= x.size)):
        raise ValueError("Downsampling factor must verify: 1 < f <= size")
    n = int(np.floor(x.size / factor))
    m = n // bins
    d = _lib.to_device(x)
    ds = _lib.empty((max(n, 1),))
    _lib.call("rt_downsample", _lib.ptr(d), x.size, float(factor), _lib.ptr(ds), _lib.stream())
    if m < 1:
        raise ValueError("fewer samples than bins after downsampling")
    scale = float(np.float32((m * factor) ** -0.5))
    total = subints == 1 or m == 1
    out = _lib.empty((bins,))
    _lib.call("rt_fold_rows", _lib.ptr(ds), m, bins, scale, 1 if total else 0, _lib.ptr(out),
              _lib.stream())
    if total:
        return _lib.to_numpy(out)
    folded = ds[:m * bins].reshape(m, bins)
    if subints is None or subints == m:
        return _lib.to_numpy(folded.contiguous())
    vf = m / subints
    if not vf < m:
        raise ValueError("factor must be strictly smaller than the number of input lines")
    nv = int(np.floor(m / vf))
    outv = _lib.empty((max(nv, 1), bins))
    _lib.call("rt_downsample_vertical", _lib.ptr(folded), m, bins, float(vf), _lib.ptr(outv),
              _lib.stream())
    return _lib.to_numpy(outv)[:nv]
