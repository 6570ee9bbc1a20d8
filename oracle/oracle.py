"""CPU oracle for the riptide FFA periodogram path.

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of bench.py import this module, and only as the
checker (or as the timed CPU baseline).  ``riptide_amd`` never imports it.

Two native libraries are wrapped:

* ``liboracle.so``: this directory's plain-C restatement of the reference
  kernels (``riptide_oracle.c``), always built (gcc) and shipped with the tree.
* ``_ref/libriptide_ref.so``: the reference's own header-only C++ compiled
  from its sources by ``Makefile`` (target ``ref``); present only where the
  reference tree was available at build time.

The numpy functions below restate the reference's Python-side arithmetic on
the path (dereddening and normalisation), which the reference itself runs in
numpy; they call the same numpy routines in the same order.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libriptide_ref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t


class OracleError(ValueError):
    pass


def _bind(lib, prefix):
    def fn(name, restype, *argtypes):
        f = getattr(lib, prefix + name)
        f.restype = restype
        f.argtypes = list(argtypes)
        return f

    b = {}
    b["last_error"] = fn("last_error", ctypes.c_char_p)
    b["ffa2"] = fn("ffa2", ctypes.c_int, _f32p, _sz, _sz, _f32p, _f32p)
    b["snr2"] = fn("snr2", ctypes.c_int, _f32p, _sz, _sz, _u64p, _sz, ctypes.c_float, _f32p)
    b["circular_prefix_sum"] = fn("circular_prefix_sum", None, _f32p, _sz, _sz, _f32p)
    b["downsample"] = fn("downsample", ctypes.c_int, _f32p, _sz, ctypes.c_double, _f32p)
    b["downsampled_variance"] = fn("downsampled_variance", ctypes.c_double, _sz, ctypes.c_double)
    b["periodogram_length"] = fn(
        "periodogram_length", ctypes.c_int, _sz, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, _sz, _sz, ctypes.POINTER(ctypes.c_size_t))
    b["periodogram"] = fn(
        "periodogram", ctypes.c_int, _f32p, _sz, ctypes.c_double, _u64p, _sz,
        ctypes.c_double, ctypes.c_double, _sz, _sz, _f64p, _u32p, _f32p)
    b["running_median"] = fn("running_median", ctypes.c_int, _f32p, _sz, _sz, _f32p)
    if prefix == "or_":
        b["periodogram_f64"] = fn(
            "periodogram_f64", ctypes.c_int, _f32p, _sz, ctypes.c_double, _u64p, _sz,
            ctypes.c_double, ctypes.c_double, _sz, _sz, _f64p)
    return b


class CLib:
    """numpy-level wrapper over one of the two native oracle libraries."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.b = _bind(self.lib, prefix)

    def _check(self, status):
        if status != 0:
            raise OracleError(self.b["last_error"]().decode())

    def ffa2(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        m, p = x.shape
        out = np.empty_like(x)
        tmp = np.empty_like(x)
        self._check(self.b["ffa2"](x, m, p, tmp, out))
        return out

    def snr2(self, x, widths, stdnoise=1.0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        widths = np.ascontiguousarray(widths, dtype=np.uint64)
        rows, cols = x.shape
        out = np.empty((rows, widths.size), dtype=np.float32)
        self._check(self.b["snr2"](x, rows, cols, widths, widths.size, stdnoise, out))
        return out

    def circular_prefix_sum(self, x, nsum):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(nsum, dtype=np.float32)
        self.b["circular_prefix_sum"](x, x.size, nsum, out)
        return out

    def downsample(self, x, f):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = int(np.floor(x.size / f)) if f > 1.0 and f <= x.size else 0
        out = np.empty(max(n, 1), dtype=np.float32)
        self._check(self.b["downsample"](x, x.size, f, out))
        return out[:n]

    def downsampled_variance(self, n, f):
        return self.b["downsampled_variance"](n, f)

    def periodogram_length(self, size, tsamp, pmin, pmax, bmin, bmax):
        L = ctypes.c_size_t(0)
        self._check(self.b["periodogram_length"](size, tsamp, pmin, pmax, bmin, bmax, ctypes.byref(L)))
        return L.value

    def periodogram(self, data, tsamp, widths, pmin, pmax, bmin, bmax):
        data = np.ascontiguousarray(data, dtype=np.float32)
        widths = np.ascontiguousarray(widths, dtype=np.uint64)
        L = self.periodogram_length(data.size, tsamp, pmin, pmax, bmin, bmax)
        periods = np.empty(L, dtype=np.float64)
        foldbins = np.empty(L, dtype=np.uint32)
        snrs = np.empty((L, widths.size), dtype=np.float32)
        self._check(self.b["periodogram"](
            data, data.size, tsamp, widths, widths.size, pmin, pmax, bmin, bmax,
            periods, foldbins, snrs))
        return periods, foldbins, snrs

    def periodogram_f64(self, data, tsamp, widths, pmin, pmax, bmin, bmax):
        """The periodogram S/N recomputed in float64 (C restatement only)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        widths = np.ascontiguousarray(widths, dtype=np.uint64)
        L = self.periodogram_length(data.size, tsamp, pmin, pmax, bmin, bmax)
        snrs = np.empty((L, widths.size), dtype=np.float64)
        self._check(self.b["periodogram_f64"](data, data.size, tsamp, widths, widths.size,
                                              pmin, pmax, bmin, bmax, snrs))
        return snrs

    def running_median(self, x, width):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        self._check(self.b["running_median"](x, x.size, width, out))
        return out


_oracle = None
_ref = None


def oracle():
    """The C restatement (always available once oracle/ is built)."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        _oracle = CLib(ORACLE_SO, "or_")
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference (only where oracle/_ref was built)."""
    global _ref
    if _ref is None:
        _ref = CLib(REF_SO, "ref_")
    return _ref


# ---------------------------------------------------------------------------
# numpy restatements of the Python-side path arithmetic
# ---------------------------------------------------------------------------

def generate_width_trials(nbins, ducy_max=0.20, wtsp=1.5):
    """ffautils.py:4-11"""
    widths = []
    w = 1
    wmax = int(max(1, ducy_max * nbins))
    while w <= wmax:
        widths.append(w)
        w = int(max(w + 1, wtsp * w))
    return np.asarray(widths)


def scrunch(data, factor):
    """running_medians.py:40-46: float32 block means (numpy pairwise sum)."""
    factor = int(factor)
    N = (data.size // factor) * factor
    return data[:N].reshape(-1, factor).mean(axis=1)


def fast_running_median(data, width_samples, min_points=101):
    """running_medians.py:49-85"""
    if not (min_points % 2):
        raise ValueError("min_points must be an odd number")
    scrunch_factor = int(max(1, width_samples / float(min_points)))
    if scrunch_factor == 1:
        return oracle().running_median(data, width_samples)
    scrunched = scrunch(data, scrunch_factor)
    rmed_lores = oracle().running_median(scrunched, min_points)
    x_lores = np.arange(scrunched.size) * scrunch_factor + 0.5 * (scrunch_factor - 1)
    return np.interp(np.arange(data.size), x_lores, rmed_lores)


def deredden(data, tsamp, width, minpts=101):
    """time_series.py:95-124: float32 data minus the approximate running
    median, computed in float64 and stored as float32."""
    data = np.asarray(data, dtype=np.float32)
    width_samples = int(round(width / tsamp))
    rmed = fast_running_median(data, width_samples, minpts)
    return np.asarray(data - rmed, dtype=np.float32)


def normalise(data):
    """time_series.py:67-93: float64 mean/var, result stored as float32."""
    data = np.asarray(data, dtype=np.float32)
    m = data.mean(dtype=np.float64)
    v = data.var(dtype=np.float64)
    norm = v ** 0.5
    return np.asarray((data - m) / norm, dtype=np.float32)


def ffafreq(N, p, dt=1.0):
    """libffa.py:129-169 (closed form without the argument checks)."""
    f0 = 1.0 / p
    m = N // p
    if m == 1:
        f = np.asarray([f0])
    else:
        s = np.arange(m)
        f = f0 - s / (m - 1.0) * f0 ** 2
    return f / dt


def ffa_search(data, tsamp, period_min=1.0, period_max=30.0, bins_min=240, bins_max=260,
               ducy_max=0.20, wtsp=1.5, deredden_=True, rmed_width=4.0, rmed_minpts=101,
               already_normalised=False, lib=None):
    """search.py:10-93 on plain arrays: returns (searched data, widths,
    periods, foldbins, snrs)."""
    lib = lib or oracle()
    x = np.asarray(data, dtype=np.float32)
    if deredden_:
        x = deredden(x, tsamp, rmed_width, rmed_minpts)
    if not already_normalised:
        x = normalise(x)
    widths = generate_width_trials(bins_min, ducy_max=ducy_max, wtsp=wtsp)
    periods, foldbins, snrs = lib.periodogram(x, tsamp, widths, period_min, period_max,
                                              bins_min, bins_max)
    return x, widths, periods, foldbins, snrs
