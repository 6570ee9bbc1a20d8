"""Device peak detection (rt_segment_select / rt_peak_candidates + host
interpolation, fit and clustering) against the numpy oracle restatement of
the reference's find_peaks: identical peaks and threshold polynomials."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,p,nw", [(1, 1, 1), (7, 1237, 3), (3, 5000, 2), (50, 33, 12)])
def test_segment_select_exact(m, p, nw):
    import torch
    from riptide_amd import _lib
    rng = np.random.default_rng(m * p + nw)
    x = rng.standard_normal((m * p, nw)).astype(np.float32)
    x[::7] = np.round(x[::7], 1)      # ties
    x[::11] = -x[::11]
    ranks = np.array([0, p // 3, p // 2, max(0, p - 2), p - 1], np.int32)
    out = torch.empty((m, nw, len(ranks)), dtype=torch.float32, device=_lib.device())
    _lib.call("rt_segment_select", _lib.ptr(_lib.to_device(x)), nw, m, p, _lib.host_ptr(ranks),
              len(ranks), _lib.ptr(out), _lib.stream())
    got = _lib.to_numpy(out)
    srt = np.sort(x.reshape(m, p, nw), axis=1)
    for t, r in enumerate(ranks):
        assert np.array_equal(got[:, :, t], srt[:, r, :])


def _compare(periods, foldbins, widths, snrs, tobs, **kw):
    from oracle import peaks as OP
    from riptide_amd import _lib
    from riptide_amd.peak_detection import PeakFinder
    want, wpoly = OP.find_peaks(periods, foldbins, widths, snrs, tobs, **kw)
    segwidth = kw.pop("segwidth", 5.0)
    finder = PeakFinder(periods, foldbins, widths, tobs, segwidth=segwidth)
    got, gpoly = finder.find(_lib.to_device(snrs), **kw)
    assert [tuple(p) for p in got] == want
    assert set(gpoly) == set(wpoly)
    for iw in wpoly:
        assert np.array_equal(np.asarray(gpoly[iw]), np.asarray(wpoly[iw]))
    return got


def test_find_peaks_planted():
    rng = np.random.default_rng(5)
    L, nw = 200000, 4
    periods = np.sort(rng.uniform(0.3, 3.0, L))[::-1].copy()
    snrs = (rng.standard_normal((L, nw)) * 1.2 + 4.0).astype(np.float32)
    for ip in (100, 77777, 150001, 199990):
        snrs[ip - 3:ip + 3, :] += 12.0
    foldbins = rng.integers(240, 261, L).astype(np.uint32)
    got = _compare(periods, foldbins, np.array([1, 2, 4, 9]), snrs, 1000.0)
    assert len(got) >= 4
    _compare(periods, foldbins, np.array([1, 2, 4, 9]), snrs, 1000.0, smin=7.5, nstd=4.0,
             polydeg=3, clrad=0.3)
    # fewer segments than minseg: constant threshold smin
    _compare(periods, foldbins, np.array([1, 2, 4, 9]), snrs, 1000.0, smin=9.0, minseg=100000)


def test_find_peaks_on_gpu_periodogram():
    """A pulsar series searched on the GPU (ffa_search), peaks found on the
    device and by the oracle from the same S/N."""
    import riptide_amd as rt
    from riptide_amd import _lib
    from riptide_amd.peak_detection import find_peaks
    n, tsamp, p0 = 2 ** 20, 256e-6, 1.3711
    t = np.arange(n) * tsamp
    ph = (t / p0) % 1.0
    sig = np.exp(-0.5 * ((ph - 0.5) / (0.02 / 2.355)) ** 2)
    sig *= 20.0 / np.sqrt((sig ** 2).sum())
    x = (np.random.default_rng(4).standard_normal(n) + sig).astype(np.float32)
    ts = rt.TimeSeries.from_numpy_array(x, tsamp)
    _, pg = rt.ffa_search(ts, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                          rmed_width=4.0)
    snrs = np.asarray(pg.snrs, np.float32)
    got = _compare(pg.periods, pg.foldbins, pg.widths, snrs, pg.tobs)
    assert abs(got[0].period / p0 - 1) < 2e-3
    peaks, polycos = find_peaks(pg)
    assert [tuple(p) for p in peaks] == [tuple(p) for p in got]


def test_find_peaks_high_degree_threshold():
    """Threshold polynomials of degree 9 and 11 (beyond the device
    evaluation, which takes up to degree 8): the device pass applies smin
    only and the host decides the threshold in float64; peaks and
    polynomials are the oracle's (np.polyfit accepts any degree,
    peak_detection.py:111-115)."""
    rng = np.random.default_rng(15)
    L, nw = 120000, 3
    periods = np.sort(rng.uniform(0.3, 3.0, L))[::-1].copy()
    snrs = (rng.standard_normal((L, nw)) * 1.2 + 4.0).astype(np.float32)
    for ip in (500, 60000, 119000):
        snrs[ip - 3:ip + 3, :] += 11.0
    foldbins = rng.integers(240, 261, L).astype(np.uint32)
    for deg in (8, 9, 11):
        _compare(periods, foldbins, np.array([1, 3, 7]), snrs, 800.0, polydeg=deg)
