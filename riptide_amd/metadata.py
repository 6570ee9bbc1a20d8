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

    def to_dict(self):
        return dict(self)

    @classmethod
    def from_dict(cls, items):
        return cls(items)

    def __str__(self):
        return "Metadata %s" % pprint.pformat(dict(self))

    def __repr__(self):
        return str(self)
