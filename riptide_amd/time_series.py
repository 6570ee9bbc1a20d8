"""TimeSeries container (time_series.py:16-413 of the reference).

normalise, deredden and downsample run on the GPU.  The file readers
(PRESTO / SIGPROC) and candidate folding are not part of this package.
"""
import copy

import numpy as np

from . import _lib
from .libffa import downsample, generate_signal
from .metadata import Metadata
from .timing import timing


class TimeSeries(object):
    """float32 time series + sampling time (s) + Metadata."""

    def __init__(self, data, tsamp, metadata=None, copy=False):
        if copy:
            self._data = np.asarray(data, dtype=np.float32).copy()
        else:
            self._data = np.asarray(data, dtype=np.float32)
        self._dev = None
        self._tsamp = float(tsamp)
        self.metadata = Metadata(metadata) if metadata is not None else Metadata({})
        # time_series.py:49-51
        self.metadata["tobs"] = self.length

    @classmethod
    def from_device(cls, tensor, tsamp, metadata=None):
        """A TimeSeries whose samples stay in a float32 device tensor until
        ``data`` is first read."""
        ts = cls.__new__(cls)
        ts._data = None
        ts._dev = tensor
        ts._tsamp = float(tsamp)
        ts.metadata = Metadata(metadata) if metadata is not None else Metadata({})
        ts.metadata["tobs"] = ts.length
        return ts

    @property
    def data(self):
        if self._data is None:
            self._data = self._dev.cpu().numpy()
        return self._data

    @property
    def data_device(self):
        """The device tensor of the samples while the host copy has not been
        made (None otherwise)."""
        return self._dev if self._data is None else None

    @property
    def tsamp(self):
        return self._tsamp

    def copy(self):
        return copy.deepcopy(self)

    def normalise(self, inplace=False):
        """Zero mean, unit variance (time_series.py:67-93)."""
        out = normalise_array(self.data)
        if inplace:
            self._data = out
        else:
            return TimeSeries(out, self.tsamp, metadata=self.metadata)

    @timing
    def deredden(self, width, minpts=101, inplace=False):
        """Subtract an approximate running median of `width` seconds
        (time_series.py:95-124)."""
        width_samples = int(round(width / self.tsamp))
        out = deredden_array(self.data, width_samples, minpts)
        if inplace:
            self._data = out
        else:
            return TimeSeries(out, self.tsamp, metadata=self.metadata)

    def downsample(self, factor, inplace=False):
        """Downsample by a real-valued factor (time_series.py:126-151)."""
        if inplace:
            self._data = downsample(self.data, factor)
            self._tsamp *= factor
        else:
            return TimeSeries(downsample(self.data, factor), factor * self.tsamp,
                              metadata=self.metadata)

    @classmethod
    def from_presto_inf(cls, fname):
        """TimeSeries from a PRESTO .inf file and its .dat file
        (time_series.py:301-333)."""
        import warnings
        from .reading.presto import PrestoInf
        inf = PrestoInf(fname)
        ts = cls(inf.load_data(), tsamp=inf["tsamp"], metadata=Metadata.from_presto_inf(inf))
        band = ts.metadata["em_band"]
        if band in ("X-ray", "Gamma"):
            warnings.warn(
                f" You have loaded file {fname!r}, which contains data observed at a "
                f"high-energy band {band!r}. riptide is NOT designed to process low photon count "
                "time series, i.e. where the background noise statistics are non-Gaussian. Be "
                "VERY careful when interpreting any search outputs.", category=UserWarning)
        return ts

    @classmethod
    def from_sigproc(cls, fname, extra_keys=None):
        """TimeSeries from a SIGPROC dedispersed time series file, 8-bit (with
        a 'signed' header key) or 32-bit (time_series.py:335-377)."""
        from .reading.sigproc import SigprocHeader
        sig = SigprocHeader(fname, extra_keys=extra_keys)
        meta = Metadata.from_sigproc(sig)
        with open(fname, "rb") as f:
            f.seek(sig.bytesize)
            data = np.fromfile(f, dtype=sig.sample_dtype).astype(np.float32)
        return cls(data, tsamp=sig["tsamp"], metadata=meta)

    def fold(self, period, bins, subints=None):
        """Fold at `period` seconds into `bins` phase bins
        (time_series.py:153-174, folding.py:21-85)."""
        from .folding import fold
        return fold(self, period, bins, subints=subints)

    @classmethod
    def generate(cls, length, tsamp, period, phi0=0.5, ducy=0.02, amplitude=10.0,
                 stdnoise=1.0):
        """Fake pulsar: von Mises pulses + white noise (time_series.py:176-235)."""
        nsamp = int(round(length / tsamp))
        data = generate_signal(nsamp, period / tsamp, phi0=phi0, ducy=ducy,
                               amplitude=amplitude, stdnoise=stdnoise)
        metadata = Metadata({
            "source_name": "fake",
            "signal_shape": "Von Mises",
            "signal_period": period,
            "signal_initial_phase": phi0,
            "signal_duty_cycle": ducy,
        })
        return cls(data, tsamp, copy=False, metadata=metadata)

    @classmethod
    def from_numpy_array(cls, array, tsamp, copy=False):
        return cls(array, tsamp, copy=copy)

    @classmethod
    def from_binary(cls, fname, tsamp, dtype=np.float32):
        return cls(np.fromfile(fname, dtype=dtype), tsamp, copy=False)

    @classmethod
    def from_npy_file(cls, fname, tsamp):
        return cls(np.load(fname), tsamp, copy=False)

    @property
    def nsamp(self):
        return self._data.size if self._data is not None else int(self._dev.numel())

    @property
    def length(self):
        return self.nsamp * self.tsamp

    @property
    def tobs(self):
        return self.length

    def __str__(self):
        return "{name} {{nsamp = {x.nsamp:d}, tsamp = {x.tsamp:.4e}, tobs = {x.length:.3f}}}".format(
            name=type(self).__name__, x=self)

    def __repr__(self):
        return str(self)

    @classmethod
    def from_dict(cls, items):
        return cls(items["data"], items["tsamp"], metadata=items["metadata"], copy=False)

    def to_dict(self):
        return {"data": self.data, "tsamp": self.tsamp, "metadata": self.metadata}


def normalise_array(data):
    x = np.ascontiguousarray(data, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty((x.size,))
    _lib.call("rt_normalise", _lib.ptr(d), x.size, _lib.ptr(out), _lib.stream())
    return _lib.to_numpy(out)


def deredden_array(data, width_samples, minpts=101):
    x = np.ascontiguousarray(data, dtype=np.float32)
    d = _lib.to_device(x)
    out = _lib.empty((x.size,))
    _lib.call("rt_deredden", _lib.ptr(d), x.size, int(width_samples), int(minpts), _lib.ptr(out),
              _lib.stream())
    return _lib.to_numpy(out)
