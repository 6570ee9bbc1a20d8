"""ffa_search (search.py:10-93 of the reference): dereddening, normalisation
and the periodogram run back to back in HBM; only the results come back."""
import numpy as np

from . import _lib, libcpp
from .ffautils import generate_width_trials
from .periodogram import Periodogram
from .time_series import TimeSeries
from .timing import timing


@timing
def ffa_search(tseries, period_min=1.0, period_max=30.0, fpmin=8, bins_min=240, bins_max=260,
               ducy_max=0.20, wtsp=1.5, deredden=True, rmed_width=4.0, rmed_minpts=101,
               already_normalised=False):
    """FFA search of one TimeSeries; returns (searched TimeSeries, Periodogram).

    `fpmin` is accepted and ignored, as in the reference (search.py:15).  With
    deredden=False and already_normalised=True the input object itself is
    returned (search.py:82-86)."""
    d = None
    n = tseries.nsamp
    if deredden or not already_normalised:
        d = _lib.to_device(tseries.data)
    if deredden:
        width_samples = int(round(rmed_width / tseries.tsamp))
        out = _lib.empty((n,))
        _lib.call("rt_deredden", _lib.ptr(d), n, width_samples, int(rmed_minpts), _lib.ptr(out),
                  _lib.stream())
        d = out
        tseries = TimeSeries(_lib.to_numpy(d), tseries.tsamp, metadata=tseries.metadata)
    if not already_normalised:
        out = _lib.empty((n,))
        _lib.call("rt_normalise", _lib.ptr(d), n, _lib.ptr(out), _lib.stream())
        d = out
        tseries = TimeSeries(_lib.to_numpy(d), tseries.tsamp, metadata=tseries.metadata)
    if d is None:
        d = _lib.to_device(tseries.data)

    widths = generate_width_trials(bins_min, ducy_max=ducy_max, wtsp=wtsp)
    periods, foldbins, snrs = libcpp.periodogram_device(
        d, n, tseries.tsamp, widths, period_min, period_max, bins_min, bins_max)
    pgram = Periodogram(widths, _lib.to_numpy(periods),
                        _lib.to_numpy(foldbins).view(np.uint32), _lib.to_numpy(snrs),
                        metadata=tseries.metadata)
    return tseries, pgram
