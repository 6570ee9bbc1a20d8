"""The GPU batch driver (staging -> BatchSearch -> device find_peaks per
range configuration) against the per-file path (from_sigproc -> ffa_search
-> find_peaks), on SIGPROC files holding pulsars at known periods."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_tim(path, data, tsamp, dm):
    def s(x):
        return struct.pack("i", len(x)) + x.encode()
    hdr = (s("HEADER_START") + s("nchans") + struct.pack("i", 1) + s("nbits") +
           struct.pack("i", 32) + s("tsamp") + struct.pack("d", tsamp) + s("refdm") +
           struct.pack("d", dm) + s("tstart") + struct.pack("d", 60000.0) + s("src_raj") +
           struct.pack("d", 120000.0) + s("src_dej") + struct.pack("d", 100000.0) +
           s("HEADER_END"))
    with open(path, "wb") as f:
        f.write(hdr + data.astype(np.float32).tobytes())


def test_batch_driver_matches_per_file_search(tmp_path):
    import riptide_amd as rt
    from riptide_amd.batch_driver import BatchWorker
    rng = np.random.default_rng(8)
    tsamp = 256e-6
    files, periods = [], [0.61, 0.73, 0.97, 1.21, 1.55, 1.83]
    for i, p0 in enumerate(periods):
        n = int(2 ** 18 + 1024 * i)
        t = np.arange(n) * tsamp
        ph = (t / p0) % 1.0
        sig = np.exp(-0.5 * ((ph - 0.5) / 0.01) ** 2)
        sig *= 18.0 / np.sqrt((sig ** 2).sum())
        x = rng.standard_normal(n) + np.cumsum(rng.standard_normal(n)) * 0.002 + sig
        f = str(tmp_path / f"dm{i}.tim")
        _write_tim(f, x, tsamp, 10.0 * i)
        files.append(f)
    search = {"period_min": 0.5, "period_max": 2.0, "bins_min": 240, "bins_max": 260}
    peaks_kw = {"smin": 7.0, "nstd": 6.0, "clrad": 0.2}
    worker = BatchWorker({"rmed_width": 1.0, "rmed_minpts": 101},
                         [{"ffa_search": search, "find_peaks": peaks_kw}], batch_size=4,
                         fmt="sigproc")
    got = worker.process_fname_list(files)
    want = []
    for f in files:
        ts = rt.TimeSeries.from_sigproc(f)
        _, pg = rt.ffa_search(ts, rmed_width=1.0, **search)
        pk, _ = rt.find_peaks(pg, **peaks_kw)
        want.extend(pk)
    key = lambda p: (p.dm, p.iw, p.ip)
    assert sorted(map(key, got)) == sorted(map(key, want))
    w = {key(p): p for p in want}
    for p in got:
        assert abs(p.snr - w[key(p)].snr) <= 1e-5 * abs(p.snr) + 1e-4
    # every pulsar is found at its period
    for i, p0 in enumerate(periods):
        best = max((p for p in got if p.dm == 10.0 * i), key=lambda p: p.snr)
        assert abs(best.period / p0 - 1) < 3e-3


def test_batch_driver_keeps_batch_across_chunks(tmp_path):
    """Chunks of series with equal lengths and sampling times reuse one kept
    batch (rebound to each chunk's series); the peaks equal the per-file
    search's."""
    import riptide_amd as rt
    from riptide_amd.batch_driver import BatchWorker
    rng = np.random.default_rng(11)
    tsamp, n = 256e-6, 2 ** 18
    files, periods = [], [0.67, 0.81, 1.09, 1.37, 1.63, 1.91]
    for i, p0 in enumerate(periods):
        t = np.arange(n) * tsamp
        ph = (t / p0) % 1.0
        sig = np.exp(-0.5 * ((ph - 0.5) / 0.01) ** 2)
        sig *= 16.0 / np.sqrt((sig ** 2).sum())
        x = rng.standard_normal(n) + sig
        f = str(tmp_path / f"dm{i}.tim")
        _write_tim(f, x, tsamp, 5.0 * i)
        files.append(f)
    search = {"period_min": 0.5, "period_max": 2.0, "bins_min": 240, "bins_max": 260}
    peaks_kw = {"smin": 7.0, "nstd": 6.0, "clrad": 0.2}
    worker = BatchWorker({"rmed_width": 1.0, "rmed_minpts": 101},
                         [{"ffa_search": search, "find_peaks": peaks_kw}], batch_size=2,
                         fmt="sigproc")
    try:
        got = worker.process_chunk(files[:2])
        kept = worker._kept[0][1]
        for i in range(2, len(files), 2):
            got.extend(worker.process_chunk(files[i:i + 2]))
        assert worker._kept[0][1] is kept  # the same batch served every chunk
    finally:
        worker.close()
    assert not worker._kept
    want = []
    for f in files:
        ts = rt.TimeSeries.from_sigproc(f)
        _, pg = rt.ffa_search(ts, rmed_width=1.0, **search)
        pk, _ = rt.find_peaks(pg, **peaks_kw)
        want.extend(pk)
    key = lambda p: (p.dm, p.iw, p.ip)
    assert sorted(map(key, got)) == sorted(map(key, want))
    w = {key(p): p for p in want}
    for p in got:
        assert abs(p.snr - w[key(p)].snr) <= 1e-5 * abs(p.snr) + 1e-4
    for i, p0 in enumerate(periods):
        best = max((p for p in got if p.dm == 5.0 * i), key=lambda p: p.snr)
        assert abs(best.period / p0 - 1) < 3e-3
