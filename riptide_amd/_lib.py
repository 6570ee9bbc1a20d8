"""ctypes binding of the C-ABI library ``libriptide_amd.so`` (include/riptide_amd.h).

Device memory and streams come from PyTorch-ROCm (plumbing only): device
buffers are torch tensors whose ``data_ptr()`` is handed to the C-ABI together
with the current HIP stream.  There is no CPU fallback: without the built
library or without a HIP device every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libriptide_amd.so")

RT_OK, RT_EINVAL, RT_EDEVICE = 0, 1, 2

_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_c_flt = ctypes.c_float
_c_ptr = ctypes.c_void_p

# name -> argtypes (all return int status)
_SIGNATURES = {
    "rt_version": [],
    "rt_set_ffa_shape": [ctypes.c_int],
    "rt_release_scratch": [],
    "rt_periodogram_length": [_c_i64, _c_dbl, _c_dbl, _c_dbl, _c_i64, _c_i64, _c_ptr],
    "rt_periodogram_grid": [_c_i64, _c_dbl, _c_dbl, _c_dbl, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_downsample_index": [_c_i64, _c_dbl, _c_i64, _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_ffa2": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_snr2": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_i64, _c_flt, _c_ptr, _c_ptr],
    "rt_downsample": [_c_ptr, _c_i64, _c_dbl, _c_ptr, _c_ptr],
    "rt_periodogram": [_c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_i64, _c_i64,
                       _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_periodogram_plain": [_c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_i64,
                             _c_i64, _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_periodogram_noalloc": [_c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_i64,
                               _c_i64, _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_running_median": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_rollback": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_fused_rollback_add": [_c_ptr, _c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_circular_prefix_sum": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_benchmark_ffa2": [_c_i64, _c_i64, _c_i64, _c_ptr],
    "rt_time_ffa": [_c_i64, _c_i64, ctypes.c_int, ctypes.c_int, _c_i64, _c_ptr],
    "rt_scrunch_mean": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_deredden": [_c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_fast_running_median": [_c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_normalise": [_c_ptr, _c_i64, _c_ptr, _c_ptr],
    "rt_normalise_ex": [_c_ptr, _c_i64, _c_ptr, _c_ptr, _c_ptr],
    "rt_normalise_backward": [_c_ptr, _c_ptr, _c_i64, _c_flt, _c_ptr, _c_ptr],
    "rt_ffa2_backward": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr],
    "rt_snr2_backward": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_i64, _c_flt, _c_ptr, _c_ptr, _c_ptr],
    "rt_periodogram_backward": [_c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_i64,
                                _c_i64, _c_ptr, _c_ptr, _c_ptr],
    "rt_batch_create": [_c_i64, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_dbl, _c_dbl, _c_ptr, _c_ptr,
                        ctypes.c_int, _c_dbl, _c_i64, ctypes.c_int, _c_i64, _c_ptr],
    "rt_batch_length": [_c_ptr, _c_i64, _c_ptr],
    "rt_batch_grid": [_c_ptr, _c_i64, _c_ptr, _c_ptr],
    "rt_batch_bind": [_c_ptr, _c_ptr, _c_ptr],
    "rt_batch_run": [_c_ptr, ctypes.c_int, _c_ptr],
    "rt_batch_run_timed": [_c_ptr, _c_ptr, _c_ptr],
    "rt_batch_check": [_c_ptr, _c_ptr],
    "rt_batch_stats": [_c_ptr, _c_ptr],
    "rt_batch_set_tiling": [_c_ptr, ctypes.c_int],
    "rt_batch_destroy": [_c_ptr],
    "rt_grad_batch_create": [_c_i64, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_i64,
                             _c_i64, _c_ptr],
    "rt_grad_batch_length": [_c_ptr, _c_ptr],
    "rt_grad_batch_normalise": [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_grad_batch_backward": [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr],
    "rt_grad_batch_destroy": [_c_ptr],
    "rt_fir_forward": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_i64, ctypes.c_int, _c_ptr, _c_ptr],
    "rt_fir_workspace_floats": [_c_i64, _c_i64, _c_i64, _c_ptr],
    "rt_fir_grad_taps": [_c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_ptr, _c_ptr],
    "rt_adam_update": [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_dbl, _c_dbl, _c_dbl, _c_dbl,
                       _c_i64, _c_ptr],
    "rt_convert_samples": [_c_ptr, _c_i64, ctypes.c_int, ctypes.c_int, _c_ptr, _c_ptr],
    "rt_fold_rows": [_c_ptr, _c_i64, _c_i64, _c_flt, ctypes.c_int, _c_ptr, _c_ptr],
    "rt_downsample_vertical": [_c_ptr, _c_i64, _c_i64, _c_dbl, _c_ptr, _c_ptr],
    "rt_segment_select": [_c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_i64, _c_ptr, _c_ptr],
    "rt_peak_candidates": [_c_ptr, _c_i64, _c_i64, _c_ptr, _c_ptr, _c_i64, _c_flt, _c_ptr, _c_ptr,
                           _c_ptr, _c_i64, _c_ptr],
}

_lib = None


def library_path():
    return LIB_PATH


def lib():
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"riptide_amd native library missing ({LIB_PATH}); "
                "build it with `python -m riptide_amd._build`")
        L = ctypes.CDLL(LIB_PATH)
        L.rt_last_error.restype = ctypes.c_char_p
        L.rt_last_error.argtypes = []
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return ["rt_last_error", *_SIGNATURES.keys()]


def call(name, *args):
    """Calls a C-ABI entry point and turns its status into an exception."""
    L = lib()
    st = getattr(L, name)(*args)
    if st == RT_OK:
        return
    msg = L.rt_last_error().decode()
    if st == RT_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def release_scratch():
    """Gives back the device memory the library keeps between calls on the
    current device (its scratch pool and the kept batch of the last
    one-series periodogram); later calls allocate again as they need."""
    call("rt_release_scratch")


# ---------------------------------------------------------------------------
# device plumbing (torch-ROCm)
# ---------------------------------------------------------------------------

def torch():
    import torch as _t
    return _t


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("riptide_amd needs a HIP GPU (MI355X); no device is available")
    return t.device("cuda", t.cuda.current_device())


def stream():
    """Handle of the current HIP stream, as a C pointer."""
    return _c_ptr(torch().cuda.current_stream().cuda_stream)


def stream_handle(s):
    """Handle of a given torch stream, as a C pointer."""
    return _c_ptr(s.cuda_stream)


def ptr(t):
    return _c_ptr(t.data_ptr())


def host_ptr(a):
    return _c_ptr(a.ctypes.data)


def to_device(a, dtype=None):
    """numpy array -> contiguous device tensor (float32 unless dtype given)."""
    t = torch()
    a = np.ascontiguousarray(a, dtype=dtype or np.float32)
    return t.from_numpy(a).to(device(), non_blocking=False)


def empty(shape, dtype="float32"):
    t = torch()
    dt = {"float32": t.float32, "float64": t.float64, "uint32": t.int32, "int32": t.int32}[dtype]
    return t.empty(shape, dtype=dt, device=device())


def to_numpy(t, dtype=None):
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype) if a.dtype.itemsize == np.dtype(dtype).itemsize else a.astype(dtype)
    return a
