"""PRESTO .inf header parsing (reading/presto.py:39-152 of the reference).

An .inf file is a list of "description = value" lines whose '=' sits at
column 40: twelve common lines (base name, telescope, instrument, object, RA,
Dec, observer, MJD, barycentred flag, sample count, sampling time, breaks
flag), the on/off bin pairs when there are breaks, then a block that depends
on the EM band (Radio: field of view, DM, lowest frequency, bandwidth,
channels, channel width, analyst; X-ray / Gamma: field of view, central
energy, energy band, analyst).  Sky positions are returned as plain (RA
hours, Dec degrees) pairs: this package does not depend on astropy."""
import os

import numpy as np

_COLUMN = 40
_FAKE = "None (Artificial Data Set)"


class SkyPosition:
    """Sky position: right ascension in hours, declination in degrees (a
    plain object, not a tuple, so that JSON encoding tags it)."""

    __slots__ = ("ra_hours", "dec_deg")

    def __init__(self, ra_hours, dec_deg):
        self.ra_hours = float(ra_hours)
        self.dec_deg = float(dec_deg)

    def __iter__(self):
        return iter((self.ra_hours, self.dec_deg))

    def __eq__(self, other):
        return isinstance(other, SkyPosition) and tuple(self) == tuple(other)

    def __repr__(self):
        return f"SkyPosition(ra_hours={self.ra_hours!r}, dec_deg={self.dec_deg!r})"


def _sexagesimal(text):
    """'[-]dd:mm:ss.sss' -> signed decimal value."""
    t = text.strip()
    sign = -1.0 if t.startswith("-") else 1.0
    parts = [float(v) for v in t.lstrip("+-").split(":")]
    while len(parts) < 3:
        parts.append(0.0)
    return sign * (parts[0] + parts[1] / 60.0 + parts[2] / 3600.0)


def _value(line, conv):
    if not (len(line) > _COLUMN and line[_COLUMN] == "="):
        raise ValueError(f"Expected '=' character at column {_COLUMN}")
    return conv(line[_COLUMN + 1:].strip())


def _flag(text):
    return int(text) != 0


def _pair(text):
    a, b = text.split(",")
    return int(a), int(b)


def parse_inf(text):
    """dict of the attributes of an .inf file's text."""
    lines = text.strip("\n").splitlines()
    head = [(
        "basename", str), ("telescope", str), ("instrument", str), ("source_name", str),
        ("raj", str), ("decj", str), ("observer", str), ("mjd", float),
        ("barycentered", _flag), ("nsamp", int), ("tsamp", float), ("breaks", _flag)]
    if _value(lines[1], str) == _FAKE:
        raise ValueError("Reading data generated with PRESTO's makedata is not supported")
    out = {k: _value(lines[i], conv) for i, (k, conv) in enumerate(head)}
    out["onoff_pairs"] = []
    rest = lines[len(head):]
    if out["breaks"]:
        for line in rest:
            try:
                out["onoff_pairs"].append(_value(line, _pair))
            except Exception:
                break
    rest = rest[len(out["onoff_pairs"]):]
    band = _value(rest[0], str)
    out["em_band"] = band
    if band == "Radio":
        block = [("fov_arcsec", float), ("dm", float), ("fbot", float), ("bandwidth", float),
                 ("nchan", int), ("cbw", float), ("analyst", str)]
    elif band in ("X-ray", "Gamma"):
        block = [("fov_arcsec", float), ("central_energy_kev", float),
                 ("energy_bandpass_kev", float), ("analyst", str)]
    else:
        raise ValueError(f"EM Band {band!r} not supported")
    for i, (k, conv) in enumerate(block):
        out[k] = _value(rest[1 + i], conv)
    return out


class PrestoInf(dict):
    """Attributes of a PRESTO .inf file (dict), with the path of its data."""

    def __init__(self, fname):
        self._fname = os.path.realpath(fname)
        with open(fname, "r") as f:
            super().__init__(parse_inf(f.read()))

    @property
    def fname(self):
        return self._fname

    @property
    def data_fname(self):
        """The associated .dat file."""
        return self.fname.rsplit(".", maxsplit=1)[0] + ".dat"

    @property
    def skycoord(self):
        return SkyPosition(_sexagesimal(self["raj"]), _sexagesimal(self["decj"]))

    def load_data(self):
        """The time series as a float32 array."""
        return np.fromfile(self.data_fname, dtype=np.float32)
