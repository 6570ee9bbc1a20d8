"""Running medians (running_medians.py:5-85 of the reference) on the GPU."""
import numpy as np

from . import _lib, libcpp


def running_median(x, width_samples):
    """Exact edge-padded running median; `width_samples` odd and < x.size
    (running_medians.py:5-37)."""
    return libcpp.running_median(np.ascontiguousarray(x), width_samples)


def scrunch(data, factor):
    """float32 means of consecutive blocks of `factor` samples, numpy's
    pairwise summation order (running_medians.py:40-46)."""
    factor = int(factor)
    x = np.ascontiguousarray(data, dtype=np.float32)
    nb = x.size // factor
    if nb == 0:
        return np.empty(0, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty((nb,))
    _lib.call("rt_scrunch_mean", _lib.ptr(d), x.size, factor, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def fast_running_median(data, width_samples, min_points=101):
    """Approximate running median: exact median of a scrunched copy, linearly
    interpolated back to full resolution (running_medians.py:49-85).  Returned
    as float32 (the reference returns numpy's float64 interpolation)."""
    if not (min_points % 2):
        raise ValueError("min_points must be an odd number")
    x = np.ascontiguousarray(data, dtype=np.float32)
    scrunch_factor = int(max(1, width_samples / float(min_points)))
    if scrunch_factor == 1:
        return running_median(x, width_samples)
    d = _lib.to_device(x)
    out = _lib.empty((x.size,))
    _lib.call("rt_fast_running_median", _lib.ptr(d), x.size, int(width_samples), int(min_points),
              _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)
