"""SIGPROC dedispersed time series headers (reading/sigproc.py:14-202 of the
reference).

The header is a sequence of (length-prefixed name, value) records between
the HEADER_START and HEADER_END markers; a name's value type is fixed by the
format (int: 32-bit, float: 64-bit double, str: length-prefixed, bool: one
unsigned byte).  The table below lists the names this reader knows; callers
can add others with `extra_keys`."""
import os
import struct

import numpy as np

from .presto import SkyPosition

KEY_TYPES = {
    **dict.fromkeys(("filename", "telescope", "rawdatafile", "source_name", "obs_date",
                     "obs_time"), str),
    **dict.fromkeys(("telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric",
                     "nbits", "nsamples", "nchans", "nifs", "nbeams", "ibeam", "hdrlen",
                     "orig_hdrlen", "new_hdrlen", "sampsize"), int),
    **dict.fromkeys(("az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1",
                     "foff", "fchannel", "refdm", "flux", "period", "pb", "ecc", "asini",
                     "bandwidth", "fbottom", "ftop", "accel"), float),
    "signed": bool,
}

_START, _END = "HEADER_START", "HEADER_END"


def _string(f):
    (n,) = struct.unpack("i", f.read(4))
    return f.read(n).decode()


def _record(f, types):
    key = _string(f)
    if key == _END:
        return key, None
    t = types.get(key)
    if t is None:
        raise KeyError(f"Type of SIGPROC header attribute '{key}' is unknown, please specify it")
    if t is str:
        return key, _string(f)
    if t is int:
        return key, struct.unpack("i", f.read(4))[0]
    if t is float:
        return key, struct.unpack("d", f.read(8))[0]
    if t is bool:
        return key, bool(struct.unpack("B", f.read(1))[0])
    raise ValueError(f"Key '{key}' has unsupported type '{t}'")


def read_header(f, extra_keys=None):
    """(attributes, header size in bytes) of an open SIGPROC file."""
    types = dict(KEY_TYPES)
    if extra_keys:
        types.update(extra_keys)
    f.seek(0)
    flag = _string(f)
    if flag != _START:
        raise ValueError(f"File starts with '{flag}' flag instead of the expected '{_START}'")
    attrs = {}
    while True:
        key, val = _record(f, types)
        if key == _END:
            return attrs, f.tell()
        attrs[key] = val


def _float_coord(v):
    """SIGPROC's ddmmss.s / hhmmss.s coordinate -> decimal hours or degrees."""
    sign = float(np.sign(v))
    x = abs(v)
    hh, x = divmod(x, 10000.0)
    mm, ss = divmod(x, 100.0)
    return sign * (hh + mm / 60.0 + ss / 3600.0)


class SigprocHeader(dict):
    """Attributes of a SIGPROC file's header (dict)."""

    def __init__(self, fname, extra_keys=None):
        self._fname = os.path.abspath(fname)
        with open(self._fname, "rb") as f:
            attrs, self._bytesize = read_header(f, extra_keys)
        super().__init__(attrs)

    @property
    def fname(self):
        return self._fname

    @property
    def bytesize(self):
        return self._bytesize

    @property
    def bytes_per_sample(self):
        return self["nchans"] * self["nbits"] // 8

    @property
    def nsamp(self):
        return (os.path.getsize(self.fname) - self.bytesize) // self.bytes_per_sample

    @property
    def tobs(self):
        return self.nsamp * self["tsamp"]

    @property
    def skycoord(self):
        return SkyPosition(_float_coord(self["src_raj"]), _float_coord(self["src_dej"]))

    def check_time_series(self):
        """Raises unless the file holds a one-channel 8-bit (with a 'signed'
        key) or 32-bit time series (metadata.py:88-103)."""
        if self["nchans"] > 1:
            raise ValueError(
                f"File {self.fname!r} contains multi-channel data (nchans = {self['nchans']}), "
                "instead of a dedispersed time series")
        nbits = self["nbits"]
        if nbits not in (8, 32):
            raise ValueError("Only 8-bit and 32-bit SIGPROC data are supported. "
                             f"File {self.fname!r} contains {nbits}-bit data")
        if nbits == 8 and "signed" not in self:
            raise ValueError("SIGPROC Header says this is 8-bit data, but does not specify its "
                             "signedness via the 'signed' key")

    @property
    def sample_dtype(self):
        if self["nbits"] == 8:
            return np.int8 if self["signed"] else np.uint8
        return np.float32
