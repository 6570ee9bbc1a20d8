"""Drop-in replacement of the reference's native module ``riptide.libcpp``
(src/riptide/cpp/python_bindings.cpp:225-279), running on the GPU.

Same ten functions, same argument meaning, same conversions (float64/int
inputs are cast to float32, non-contiguous float32 input raises) and the same
ValueError messages.  numpy in, numpy out; every call stages its input in
HBM, runs the HIP kernels and copies the result back.
"""
import numpy as np

from . import _lib, tune


def _f32(arr, ndim):
    """pybind11 py::array_t<float> (forcecast) + assert_c_contiguous + unchecked<N>
    (python_bindings.cpp:33-41)."""
    a = np.asarray(arr)
    if a.dtype != np.float32:
        a = a.astype(np.float32)
    if not a.flags.c_contiguous:
        raise ValueError("Input numpy array must be contiguous in memory")
    if a.ndim != ndim:
        raise ValueError(
            f"array has incorrect number of dimensions: {a.ndim}; expected {ndim}")
    return a


def _widths(widths):
    w = np.asarray(widths)
    if w.dtype != np.uint64:
        w = w.astype(np.uint64)
    if not w.flags.c_contiguous:
        raise ValueError("Input numpy array must be contiguous in memory")
    if w.ndim != 1:
        raise ValueError(f"array has incorrect number of dimensions: {w.ndim}; expected 1")
    return w


def _size_t(v, name):
    if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)):
        raise TypeError(f"{name} must be an integer")
    if v < 0:
        raise TypeError(f"{name} must be a non-negative integer")
    return int(v)


def rollback(x, shift):
    """roll(x, -shift) (python_bindings.cpp:44-52)."""
    x = _f32(x, 1)
    shift = _size_t(shift, "shift")
    if x.size == 0:
        return np.empty(0, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty((x.size,))
    _lib.call("rt_rollback", _lib.ptr(d), x.size, shift, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def fused_rollback_add(x, y, shift):
    """x + roll(y, -shift) (python_bindings.cpp:55-70)."""
    x = np.asarray(x)
    y = np.asarray(y)
    if x.size != y.size:
        raise ValueError("Arrays must have the same number of elements")
    x = _f32(x, 1)
    y = _f32(y, 1)
    shift = _size_t(shift, "shift")
    if x.size == 0:
        return np.empty(0, dtype=np.float32)
    dx, dy = _lib.to_device(x), _lib.to_device(y)
    out = _lib.empty((x.size,))
    _lib.call("rt_fused_rollback_add", _lib.ptr(dx), _lib.ptr(dy), x.size, shift, _lib.ptr(out),
              _lib.stream())
    return _lib.to_numpy(out)


def circular_prefix_sum(x, nsum):
    """Circular prefix sum over nsum elements (python_bindings.cpp:73-81).

    Returns nsum elements; the reference allocates only x.size of them and
    overruns its buffer when nsum > x.size."""
    x = _f32(x, 1)
    nsum = _size_t(nsum, "nsum")
    if nsum == 0 or x.size == 0:
        return np.empty(nsum if x.size else 0, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty((nsum,))
    _lib.call("rt_circular_prefix_sum", _lib.ptr(d), x.size, nsum, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def ffa2(data):
    """FFA transform of a 2-D float32 array (python_bindings.cpp:84-96)."""
    x = _f32(data, 2)
    rows, cols = x.shape
    d = _lib.to_device(x)
    out = _lib.empty((rows, cols))
    _lib.call("rt_ffa2", _lib.ptr(d), rows, cols, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def benchmark_ffa2(rows, cols, loops):
    """Seconds per FFA transform of a zero rows x cols block
    (python_bindings.cpp:99-118), timed with HIP events."""
    import ctypes
    sec = ctypes.c_double(0.0)
    _lib.device()
    _lib.call("rt_benchmark_ffa2", _size_t(rows, "rows"), _size_t(cols, "cols"),
              _size_t(loops, "loops"), ctypes.byref(sec))
    return sec.value


def _snr(x, widths, stdnoise):
    w = _widths(widths)
    rows, cols = x.shape
    stdnoise = float(np.float32(stdnoise))
    out_shape = (rows, w.size)
    # argument checks in the reference's order (python_bindings.cpp:131-132:
    # check_stdnoise, then check_trial_widths, snr.hpp:14-32), on the host
    # before any device work; rt_snr2 repeats them
    if not stdnoise > 0:
        raise ValueError("stdnoise must be > 0")
    if not bool(np.all((w > 0) & (w < cols))):
        raise ValueError("trial widths must be all > 0 and < columns")
    if rows == 0 or w.size == 0:
        return np.zeros(out_shape, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty(out_shape)
    _lib.call("rt_snr2", _lib.ptr(d), rows, cols, _lib.host_ptr(w), w.size, stdnoise,
              _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def snr1(data, widths, stdnoise=1.0):
    """S/N of one profile for several boxcar widths (python_bindings.cpp:121-138)."""
    x = _f32(data, 1)
    return _snr(x.reshape(1, -1), widths, stdnoise).reshape(-1)


def snr2(data, widths, stdnoise=1.0):
    """S/N of (num_profiles, num_bins) profiles (python_bindings.cpp:141-160)."""
    x = _f32(data, 2)
    return _snr(x, widths, stdnoise)


def downsample(data, factor):
    """Downsample by a real-valued factor (python_bindings.cpp:163-177)."""
    x = _f32(data, 1)
    f = float(factor)
    if not ((f > 1.0) and (f <= x.size)):
        raise ValueError("Downsampling factor must verify: 1 < f <= size")
    n = int(np.floor(x.size / f))
    d = _lib.to_device(x)
    out = _lib.empty((max(n, 1),))
    _lib.call("rt_downsample", _lib.ptr(d), x.size, f, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)[:n]


def periodogram_length(size, tsamp, period_min, period_max, bins_min, bins_max):
    import ctypes
    L = ctypes.c_int64(0)
    _lib.call("rt_periodogram_length", int(size), float(tsamp), float(period_min),
              float(period_max), _size_t(bins_min, "bins_min"), _size_t(bins_max, "bins_max"),
              ctypes.byref(L))
    return L.value


def periodogram_device(d_data, size, tsamp, widths, period_min, period_max, bins_min, bins_max,
                       plain=False):
    """Periodogram of a series already in HBM; returns device tensors
    (periods float64, foldbins int32 holding uint32 bits, snrs float32)."""
    w = _widths(widths)
    L = periodogram_length(size, tsamp, period_min, period_max, bins_min, bins_max)
    periods = _lib.empty((L,), "float64")
    foldbins = _lib.empty((L,), "uint32")
    snrs = _lib.empty((L, w.size))
    name = "rt_periodogram_plain" if plain else "rt_periodogram"
    # FFA workgroup shape chosen by the autotuner for this bins range, if any;
    # rt_periodogram runs the search as a batch of one series, tiled as the
    # tuned (rows, bins) classes say (none while a shape is forced)
    shape = tune.shape_for("periodogram", int(bins_min))
    if shape is not None:
        _lib.call("rt_set_ffa_shape", shape)
    if not plain:
        tune.set_batch_tiling(tune.batch_tiling())
    try:
        _lib.call(name, _lib.ptr(d_data), int(size), float(tsamp), _lib.host_ptr(w), w.size,
                  float(period_min), float(period_max), int(bins_min), int(bins_max),
                  _lib.ptr(periods), _lib.ptr(foldbins), _lib.ptr(snrs), _lib.stream())
    finally:
        if not plain:
            tune.set_batch_tiling([])
        if shape is not None:
            tune.restore()
    return periods, foldbins, snrs


def periodogram(data, tsamp, widths, period_min, period_max, bins_min, bins_max):
    """Periodogram of a normalised series (python_bindings.cpp:180-209):
    returns (periods float64, foldbins uint32, snrs float32 [L, nw])."""
    x = _f32(data, 1)
    w = _widths(widths)
    bins_min = _size_t(bins_min, "bins_min")
    bins_max = _size_t(bins_max, "bins_max")
    d = _lib.to_device(x)
    periods, foldbins, snrs = periodogram_device(d, x.size, tsamp, w, period_min, period_max,
                                                 bins_min, bins_max)
    return (_lib.to_numpy(periods), _lib.to_numpy(foldbins).view(np.uint32),
            _lib.to_numpy(snrs))


def running_median(data, width):
    """Edge-padded running median (python_bindings.cpp:212-222)."""
    x = _f32(data, 1)
    width = _size_t(width, "width")
    if not (width % 2):
        raise ValueError("width must be an odd number")
    if not (width < x.size):
        raise ValueError("width must be < size")
    d = _lib.to_device(x)
    out = _lib.empty((x.size,))
    _lib.call("rt_running_median", _lib.ptr(d), x.size, width, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)
