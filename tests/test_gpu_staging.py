"""Batched staging of time series files to the device (one pinned buffer,
one transfer, 8-bit samples widened on the GPU) against the host readers."""
import os
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "readers")


def _tim(path, data, nbits, signed):
    def s(x):
        return struct.pack("i", len(x)) + x.encode()
    hdr = s("HEADER_START") + s("nchans") + struct.pack("i", 1) + s("nbits") + \
        struct.pack("i", nbits) + s("tsamp") + struct.pack("d", 64e-6)
    if signed is not None:
        hdr += s("signed") + struct.pack("B", int(signed))
    hdr += s("src_raj") + struct.pack("d", 123456.7) + s("src_dej") + struct.pack("d", -101010.0)
    hdr += s("HEADER_END")
    with open(path, "wb") as f:
        f.write(hdr + data.tobytes())


def test_stage_reference_fixtures():
    import riptide_amd as rt
    from riptide_amd import _lib
    from riptide_amd.staging import stage_files
    names = ["fake_presto_radio.inf", "fake_sigproc_float32.tim", "fake_sigproc_uint8.tim",
             "fake_sigproc_int8.tim", "fake_presto_radio_breaks.inf"]
    b = stage_files([os.path.join(DATA, n) for n in names])
    for x, t in zip(b.series, b.tsamps):
        assert np.array_equal(_lib.to_numpy(x), np.arange(16, dtype=np.float32)) and t == 64e-6


def test_stage_many_files_matches_host_readers(tmp_path):
    import riptide_amd as rt
    from riptide_amd import _lib
    from riptide_amd.staging import stage_files
    rng = np.random.default_rng(0)
    files = []
    for i in range(12):
        n = int(rng.integers(1000, 300000))
        kind = i % 3
        if kind == 0:
            d, nb, sg = rng.standard_normal(n).astype(np.float32), 32, None
        elif kind == 1:
            d, nb, sg = rng.integers(0, 256, n).astype(np.uint8), 8, False
        else:
            d, nb, sg = rng.integers(-128, 128, n).astype(np.int8), 8, True
        p = str(tmp_path / f"s{i}.tim")
        _tim(p, d, nb, sg)
        files.append(p)
    b = stage_files(files)
    assert b.sizes == [rt.TimeSeries.from_sigproc(f).nsamp for f in files]
    for f, x in zip(files, b.series):
        assert np.array_equal(_lib.to_numpy(x), rt.TimeSeries.from_sigproc(f).data)
