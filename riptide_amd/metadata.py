"""Observation metadata (metadata.py:25-125 of the reference), validated in
plain Python instead of with the ``schema`` package.

Reserved keys and their rules (metadata.py:11-21): source_name str, skycoord
any sky-coordinate object, dm float >= 0, mjd float >= 0, tobs float > 0,
fname str, each possibly None; any other key must be a str whose value is
JSON-serialisable.  Missing reserved keys are set to None.
"""
import json
import pprint

RESERVED = ("source_name", "skycoord", "dm", "mjd", "tobs", "fname")


def _is_float(x):
    return isinstance(x, float)


_RULES = {
    "source_name": lambda v: v is None or isinstance(v, str),
    "skycoord": lambda v: True,
    "dm": lambda v: v is None or (_is_float(v) and v >= 0),
    "mjd": lambda v: v is None or (_is_float(v) and v >= 0),
    "tobs": lambda v: v is None or (_is_float(v) and v > 0),
    "fname": lambda v: v is None or isinstance(v, str),
}


class MetadataError(ValueError):
    pass


def validate(items):
    for k, v in dict(items).items():
        if k in _RULES:
            if not _RULES[k](v):
                raise MetadataError(f"Metadata key {k!r} has an invalid value: {v!r}")
        elif isinstance(k, str):
            try:
                json.dumps(v)
            except (TypeError, ValueError) as err:
                raise MetadataError(f"Metadata key {k!r} is not JSON-serialisable") from err
        # non-str keys are ignored by the reference's schema (ignore_extra_keys)


class Metadata(dict):
    """dict of observation attributes carried by TimeSeries and Periodogram."""

    def __init__(self, items={}):
        validate(items)
        super().__init__(items)
        for k in RESERVED:
            self.setdefault(k, None)

    @classmethod
    def from_presto_inf(cls, inf):
        """Metadata of a PRESTO .inf file (metadata.py:53-71); `inf` is a
        PrestoInf or a path."""
        import os
        from .reading.presto import PrestoInf
        if isinstance(inf, str):
            inf = PrestoInf(inf)
        attrs = dict(inf)
        attrs["skycoord"] = inf.skycoord
        attrs["fname"] = os.path.realpath(inf.fname)
        attrs["tobs"] = attrs["tsamp"] * attrs["nsamp"]
        return cls(attrs)

    @classmethod
    def from_sigproc(cls, sh, extra_keys=None):
        """Metadata of a SIGPROC time series file (metadata.py:73-111); `sh`
        is a SigprocHeader or a path."""
        import os
        from .reading.sigproc import SigprocHeader
        if isinstance(sh, str):
            sh = SigprocHeader(sh, extra_keys=extra_keys)
        sh.check_time_series()
        attrs = dict(sh)
        attrs["dm"] = attrs.get("refdm", None)
        attrs["skycoord"] = sh.skycoord
        attrs["source_name"] = attrs.get("source_name", None)
        attrs["mjd"] = attrs.get("tstart", None)
        attrs["fname"] = os.path.realpath(sh.fname)
        attrs["tobs"] = sh.tobs
        return cls(attrs)

    def to_dict(self):
        return dict(self)

    @classmethod
    def from_dict(cls, items):
        return cls(items)

    def __str__(self):
        return "Metadata %s" % pprint.pformat(dict(self))

    def __repr__(self):
        return str(self)
