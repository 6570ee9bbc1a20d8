"""Periodogram container (periodogram.py:9-104 of the reference, without the
matplotlib plotting helpers).

The S/N array may be handed over as a device tensor: it then stays in HBM
(``snrs_device``, what find_peaks reads) and is copied to the host only when
``snrs`` is first read."""
import numpy as np

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
    def snrs(self):
        """S/N as a host numpy array (copied from the device on first use)."""
        if self._snrs is None:
            self._snrs = self._snrs_dev.cpu().numpy()
        return self._snrs

    @snrs.setter
    def snrs(self, value):
        if isinstance(value, np.ndarray) or not hasattr(value, "is_cuda"):
            self._snrs, self._snrs_dev = value, None
        else:
            self._snrs, self._snrs_dev = None, value

    @property
    def snrs_device(self):
        """The device tensor of the S/N, or None when they live on the host."""
        return self._snrs_dev

    @property
    def snrs_on_host(self):
        """Whether the host copy of the S/N exists."""
        return self._snrs is not None

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
