"""ffa_search (search.py:10-93 of the reference): dereddening, normalisation
and the periodogram run back to back in HBM.  The returned series and S/N
stay in HBM and are copied to the host when first read (TimeSeries.data,
Periodogram.snrs); the trial periods and fold bins come back at once."""
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
    n = tseries.nsamp
    d = tseries.data_device if tseries.data_device is not None else _lib.to_device(tseries.data)
    if deredden:
        width_samples = int(round(rmed_width / tseries.tsamp))
        out = _lib.empty((n,))
        _lib.call("rt_deredden", _lib.ptr(d), n, width_samples, int(rmed_minpts), _lib.ptr(out),
                  _lib.stream())
        d = out
    if not already_normalised:
        out = _lib.empty((n,))
        _lib.call("rt_normalise", _lib.ptr(d), n, _lib.ptr(out), _lib.stream())
        d = out
    if deredden or not already_normalised:
        # the searched series stays in HBM; its host copy is made on first use
        tseries = TimeSeries.from_device(d, tseries.tsamp, metadata=tseries.metadata)

    widths = generate_width_trials(bins_min, ducy_max=ducy_max, wtsp=wtsp)
    periods, foldbins, snrs = libcpp.periodogram_device(
        d, n, tseries.tsamp, widths, period_min, period_max, bins_min, bins_max)
    # trial periods and fold bins are small (12 bytes per period): host
    # arrays; the S/N stay in HBM until read
    pgram = Periodogram(widths, _lib.to_numpy(periods),
                        _lib.to_numpy(foldbins).view(np.uint32), snrs,
                        metadata=tseries.metadata)
    return tseries, pgram
