"""PRESTO / SIGPROC readers against the reference's own test data files
(copied under tests/golden/readers: 16 samples 0..15 at 64 us) and their
expected behaviour (tests/test_time_series.py:16-62 of the reference)."""
import os
import warnings

import numpy as np
import pytest

import riptide_amd as rt

DATA = os.path.join(os.path.dirname(__file__), "golden", "readers")


def _check(ts):
    assert ts.nsamp == 16 and ts.tsamp == 64e-6 and ts.data.dtype == np.float32
    assert np.array_equal(ts.data, np.arange(16))


def test_presto():
    for f in ("fake_presto_radio.inf", "fake_presto_radio_breaks.inf"):
        _check(rt.TimeSeries.from_presto_inf(os.path.join(DATA, f)))
    with pytest.warns(UserWarning):
        _check(rt.TimeSeries.from_presto_inf(os.path.join(DATA, "fake_presto_xray.inf")))
    ts = rt.TimeSeries.from_presto_inf(os.path.join(DATA, "fake_presto_radio_breaks.inf"))
    assert ts.metadata["onoff_pairs"] == [(0, 14), (15, 15)]
    assert ts.metadata["dm"] == 42.42 and ts.metadata["tobs"] == 16 * 64e-6
    ra, dec = ts.metadata["skycoord"]
    assert abs(ra - 1 / 3600) < 1e-12 and abs(dec + 1 / 3600) < 1e-12


def test_sigproc():
    for f in ("fake_sigproc_float32.tim", "fake_sigproc_uint8.tim", "fake_sigproc_int8.tim"):
        ts = rt.TimeSeries.from_sigproc(os.path.join(DATA, f))
        _check(ts)
        assert ts.metadata["mjd"] is not None and ts.metadata["dm"] is not None
    with pytest.raises(ValueError):
        rt.TimeSeries.from_sigproc(os.path.join(DATA, "fake_sigproc_uint8_nosignedkey.tim"))


def test_sigproc_unknown_key_needs_a_type(tmp_path):
    import struct

    def s(x):
        return struct.pack("i", len(x)) + x.encode()
    raw = s("HEADER_START") + s("nbits") + struct.pack("i", 32) + s("mystery") + \
        struct.pack("i", 7) + s("HEADER_END")
    p = tmp_path / "x.tim"
    p.write_bytes(raw)
    from riptide_amd.reading.sigproc import SigprocHeader
    with pytest.raises(KeyError):
        SigprocHeader(str(p))
    h = SigprocHeader(str(p), extra_keys={"mystery": int})
    assert h["mystery"] == 7 and h["nbits"] == 32 and h.bytesize == len(raw)


def test_json_round_trip_and_wire_format(tmp_path):
    """save_json/load_json (serialization.py of the reference): arrays as
    base64 numpy.ndarray records, sky positions as astropy.SkyCoord
    records, objects tagged with __type__ / __version__."""
    import json
    ts = rt.TimeSeries.from_presto_inf(os.path.join(DATA, "fake_presto_radio.inf"))
    pg = rt.Periodogram(np.array([1, 2, 3]), np.linspace(1.0, 2.0, 5),
                        np.arange(240, 245, dtype=np.uint32),
                        np.arange(15, dtype=np.float32).reshape(5, 3), metadata=ts.metadata)
    for obj in (ts, pg):
        p = tmp_path / "x.json"
        rt.save_json(str(p), obj)
        raw = json.loads(p.read_text())
        assert raw["__type__"] == type(obj).__name__ and "__version__" in raw
        back = rt.load_json(str(p))
        assert type(back) is type(obj)
    assert raw["snrs"]["__type__"] == "numpy.ndarray" and raw["snrs"]["dtype"] == "float32"
    assert raw["metadata"]["skycoord"]["__type__"] == "astropy.SkyCoord"
    assert abs(raw["metadata"]["skycoord"]["rajd"] - 15.0 / 3600) < 1e-12
    assert np.array_equal(back.snrs, pg.snrs) and np.array_equal(back.periods, pg.periods)
    assert back.metadata["dm"] == 42.42
    t2 = rt.load_json(str(tmp_path / "x.json"))
    assert t2.metadata["skycoord"] == pg.metadata["skycoord"]
