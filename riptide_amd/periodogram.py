"""Periodogram container (periodogram.py:9-104 of the reference, without the
matplotlib plotting helpers)."""
from .metadata import Metadata


class Periodogram(object):
    """Raw output of the FFA search of one time series.

    widths   : trial pulse widths in phase bins (int array)
    periods  : trial periods in seconds (float64, plan order)
    foldbins : phase bins used for each trial period (uint32)
    snrs     : S/N, shape (num_periods, num_widths) (float32)
    """

    def __init__(self, widths, periods, foldbins, snrs, metadata=None):
        self.widths = widths
        self.periods = periods
        self.foldbins = foldbins
        self.snrs = snrs
        self.metadata = metadata if metadata is not None else Metadata({})

    @property
    def freqs(self):
        """Trial frequencies in Hz, in decreasing order."""
        return 1.0 / self.periods

    @property
    def tobs(self):
        """Length in seconds of the searched time series."""
        return self.metadata["tobs"]

    def to_dict(self):
        return {
            "widths": self.widths,
            "periods": self.periods,
            "foldbins": self.foldbins,
            "snrs": self.snrs,
            "metadata": self.metadata,
        }

    @classmethod
    def from_dict(cls, items):
        return cls(items["widths"], items["periods"], items["foldbins"], items["snrs"],
                   metadata=items["metadata"])
