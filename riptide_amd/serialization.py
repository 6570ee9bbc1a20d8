"""JSON serialisation of riptide objects (serialization.py:1-153 of the
reference), in the same wire format so that files written by either can be
read by the other: numpy arrays as {"__type__": "numpy.ndarray", "data":
base64 of the raw bytes, "dtype", "shape"}; sky positions as
{"__type__": "astropy.SkyCoord", "rajd", "decjd", "frame": "icrs"}; objects
with to_dict() as their dict plus "__type__" (class name) and "__version__".
Without astropy, sky positions decode to SkyPosition (RA hours, Dec
degrees); pandas DataFrames (candidate tables) are not handled."""
import base64
import json

import numpy as np

from .reading.presto import SkyPosition


def _version():
    from . import __version__
    return __version__


def _class(name):
    import riptide_amd
    cls = getattr(riptide_amd, name, None)
    if cls is None or not hasattr(cls, "from_dict"):
        raise ValueError(f"cannot decode objects of type {name!r}")
    return cls


class JSONEncoder(json.JSONEncoder):
    def default(self, obj):
        if isinstance(obj, np.ndarray):
            return {"__type__": "numpy.ndarray",
                    "data": base64.b64encode(np.ascontiguousarray(obj).data).decode(),
                    "dtype": str(obj.dtype), "shape": obj.shape}
        if isinstance(obj, np.integer):
            return int(obj)
        if isinstance(obj, np.floating):
            return float(obj)
        if isinstance(obj, SkyPosition):
            return {"__type__": "astropy.SkyCoord", "rajd": 15.0 * obj.ra_hours,
                    "decjd": obj.dec_deg, "frame": "icrs"}
        if hasattr(obj, "to_dict"):
            items = dict(obj.to_dict())
            items["__type__"] = type(obj).__name__
            items["__version__"] = getattr(obj, "version", None) or _version()
            return items
        return super().default(obj)


def object_hook(items):
    kind = items.get("__type__")
    if kind is None:
        return items
    if kind == "numpy.ndarray":
        raw = base64.b64decode(items["data"].encode())
        return np.frombuffer(raw, items["dtype"]).reshape(items["shape"])
    if kind == "astropy.SkyCoord":
        return SkyPosition(items["rajd"] / 15.0, items["decjd"])
    obj = _class(kind).from_dict(items)
    obj.version = items.get("__version__", _version())
    return obj


def to_json(obj, **kwargs):
    kwargs = dict(kwargs)
    kwargs["cls"] = JSONEncoder
    kwargs.setdefault("indent", 4)
    return json.dumps(obj, **kwargs)


def from_json(s):
    return json.loads(s, object_hook=object_hook)


def save_json(fname, obj, **kwargs):
    with open(fname, "w") as f:
        f.write(to_json(obj, **kwargs))


def load_json(fname):
    with open(fname, "r") as f:
        return from_json(f.read())
