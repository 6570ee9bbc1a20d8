"""The reference's own API tests restated against riptide_amd on the GPU
(tests/test_ffa_base_functions.py:39-57, tests/test_snr.py:25-78,
tests/test_running_median.py:33-77, tests/test_time_series.py:101-164,
tests/test_ffa_search_pgram.py:9-94 of the reference), with the expected
values taken from the reference tests' closed forms and, where they only
check shapes, from the CPU oracle on the same inputs."""
import os
import tempfile

import numpy as np
import pytest

import riptide_amd as rt

pytestmark = pytest.mark.gpu

FFA_IN_88 = np.zeros((8, 8), np.float32)
FFA_IN_88[:, 7] = 1
FFA_OUT_88 = np.array([
    [0, 0, 0, 0, 0, 0, 0, 8],
    [0, 0, 0, 0, 0, 0, 4, 4],
    [0, 0, 0, 0, 0, 2, 4, 2],
    [0, 0, 0, 0, 2, 2, 2, 2],
    [0, 0, 0, 1, 2, 2, 2, 1],
    [0, 0, 1, 2, 1, 1, 2, 1],
    [0, 1, 1, 1, 2, 1, 1, 1],
    [1, 1, 1, 1, 1, 1, 1, 1],
], np.float32)


def test_ffa1_ffa2_known_answers():
    """tests/test_ffa_base_functions.py:39-57: the 8x8 table under every
    phase roll and with 0-7 extra zero columns, through ffa2 and ffa1
    (float64 input is cast to float32)."""
    for shift in range(8):
        X = np.roll(FFA_IN_88, shift, axis=1)
        truth = np.roll(FFA_OUT_88, shift, axis=1)
        assert np.array_equal(rt.ffa2(X), truth)
        assert np.array_equal(rt.ffa1(X.ravel(), X.shape[1]), truth)
    for extra in range(8):
        X = np.hstack([FFA_IN_88, np.zeros((8, extra))])  # float64
        truth = np.hstack([FFA_OUT_88, np.zeros((8, extra), np.float32)])
        Y = rt.ffa2(X)
        assert Y.dtype == np.float32 and np.array_equal(Y, truth)
        assert np.array_equal(rt.ffa1(X.ravel(), np.int64(X.shape[1])), truth)


def test_ffa1_drops_incomplete_period(oracle):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000).astype(np.float32)
    p = 37
    m = x.size // p
    got = rt.ffa1(x, p)
    assert got.shape == (m, p)
    assert np.array_equal(got, oracle.oracle().ffa2(x[: m * p].reshape(m, p)))


def test_boxcar_snr_output_dims(oracle):
    """tests/test_snr.py:25-45, with values checked against the oracle."""
    cols, rows, layers = 32, 4, 3
    widths = [1, 2, 3, 5]
    rng = np.random.default_rng(5)
    d3 = rng.standard_normal((layers, rows, cols)).astype(np.float32)
    want = oracle.oracle().snr2(d3.reshape(-1, cols), widths)
    s1 = rt.boxcar_snr(d3[0, 0], widths)
    assert s1.ndim == 1 and s1.size == len(widths)
    s2 = rt.boxcar_snr(d3[0], widths)
    assert s2.shape == (rows, len(widths))
    s3 = rt.boxcar_snr(d3, widths)
    assert s3.shape == (layers, rows, len(widths))
    np.testing.assert_allclose(s3.reshape(-1, len(widths)), want, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(s1, want[0], rtol=1e-5, atol=1e-5)
    # snr1 of one profile equals that profile's snr2 row
    from riptide_amd import libcpp
    np.testing.assert_array_equal(libcpp.snr1(d3[1, 2], widths, 1.5),
                                  libcpp.snr2(d3[1, 2:3], widths, 1.5)[0])
    # float64 and non-contiguous inputs go through a contiguous float32 copy
    np.testing.assert_allclose(rt.boxcar_snr(d3[:, :, ::-1].astype(np.float64), widths),
                               rt.boxcar_snr(np.ascontiguousarray(d3[:, :, ::-1]), widths),
                               rtol=0, atol=0)


def test_boxcar_snr_phase_rotation_invariance():
    """tests/test_snr.py:48-59."""
    rng = np.random.default_rng(9)
    data = rng.standard_normal((4, 32)).astype(np.float32)
    widths = [1, 2, 5, 11, 18, 31]
    ref = rt.boxcar_snr(data, widths)
    for shift in range(1, 33):
        np.testing.assert_allclose(rt.boxcar_snr(np.roll(data, shift, axis=-1), widths), ref,
                                   rtol=1e-5, atol=1e-5)


def test_boxcar_snr_top_hat_values():
    """tests/test_snr.py:62-78: a top-hat of true width w peaks at trial
    width w with S/N = w*h, h = sqrt((n - w)/(n*w))."""
    n = 64
    widths = np.arange(1, n)
    data = np.zeros(n, dtype=np.float32)
    for w in range(1, n):
        data[:w] = 1.0
        snr = rt.boxcar_snr(data, widths)
        assert snr.argmax() == w - 1
        h = np.sqrt((n - w) / (n * w))
        assert np.isclose(snr.max(), w * h, rtol=1e-5)


def _naive_rmed(x, w):
    # tests/test_running_median.py:7-17 of the reference
    h = w // 2
    p = np.pad(x, (h, h), mode="edge")
    return np.asarray([np.median(p[i:i + w]) for i in range(x.size)], dtype=np.float32)


def test_running_median_naive_and_non_contiguous():
    """tests/test_running_median.py:33-52: exact against the naive median,
    also on non-contiguous column slices."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal(100).astype(np.float32)
    for w in [1, 3, 5, 7, 11, 25, 37]:
        assert np.array_equal(rt.running_median(x, w), _naive_rmed(x, w))
    data = rng.standard_normal(300).reshape(100, 3).astype(np.float32)
    for col in data.T:
        assert not col.flags.c_contiguous
        for w in [1, 3, 5, 7, 11, 25, 37]:
            assert np.array_equal(rt.running_median(col, w), _naive_rmed(col, w))


def test_fast_running_median_noscrunch():
    """tests/test_running_median.py:59-77: with min_points > width no
    scrunching happens and the result is the exact running median."""
    x = np.random.default_rng(4).standard_normal(100)
    widths = [1, 3, 5, 7, 11, 25, 37]
    wmax = max(widths)
    min_points = wmax + (wmax % 2) + 1
    for w in widths:
        assert np.array_equal(rt.running_median(x, w),
                              rt.fast_running_median(x, w, min_points=min_points))


def test_time_series_methods():
    """tests/test_time_series.py:101-164."""
    length, tsamp, period, amplitude = 10.0, 1.0e-3, 1.0, 25.0
    np.random.seed(11)
    tsorig = rt.TimeSeries.generate(length, tsamp, period, amplitude=amplitude, stdnoise=1.0)
    ts = tsorig.copy()
    tscopy = ts.copy()
    ts = ts.normalise()
    tscopy.normalise(inplace=True)
    assert np.isclose(ts.data.mean(), 0.0, atol=1e-6)
    assert np.isclose(ts.data.std(), 1.0, atol=1e-6)
    assert np.allclose(ts.data, tscopy.data, atol=1e-6)
    tscopy = ts.copy()
    ts = ts.deredden(width=0.5, minpts=51)
    tscopy.deredden(width=0.5, minpts=51, inplace=True)
    assert np.allclose(ts.data, tscopy.data, atol=1e-6)
    tsconst = rt.TimeSeries.generate(length, tsamp, period, amplitude=0, stdnoise=0)
    tsconst._data += 42.42
    assert np.allclose(tsconst.deredden(0.5, minpts=51).data, 0.0, atol=1e-6)
    dsfactor = 10
    ts = tsorig.downsample(dsfactor)
    tscopy = tsorig.copy()
    tscopy.downsample(dsfactor, inplace=True)
    for t in (ts, tscopy):
        assert t.tsamp == tsorig.tsamp * dsfactor
        assert t.nsamp == tsorig.nsamp // dsfactor
        assert t.length == tsorig.length


def test_ffa_search_and_json_round_trip():
    """tests/test_ffa_search_pgram.py:9-73 (without the matplotlib plots)."""
    np.random.seed(0)
    length, tsamp, period = 200.0, 0.001, 1.0
    ts = rt.TimeSeries.generate(length, tsamp, period, amplitude=20.0)
    kw = dict(period_min=0.8 * period, period_max=1.2 * period, bins_min=240, bins_max=260)
    tsdr, pgram = rt.ffa_search(ts, **kw)
    assert all(np.maximum.accumulate(pgram.periods) == pgram.periods)
    assert pgram.snrs.shape == (len(pgram.periods), len(pgram.widths))
    assert pgram.metadata == ts.metadata == tsdr.metadata
    assert pgram.tobs == length
    assert all(pgram.freqs == 1.0 / pgram.periods)
    # the injected period is found
    best = pgram.snrs.max(axis=1).argmax()
    assert abs(pgram.periods[best] - period) < 2e-3 and pgram.snrs.max() > 15
    tsdr2, _ = rt.ffa_search(ts, already_normalised=True, deredden=False, **kw)
    assert id(tsdr2) == id(ts)
    with tempfile.TemporaryDirectory() as d:
        fname = os.path.join(d, "pgram.json")
        rt.save_json(fname, pgram)
        back = rt.load_json(fname)
    assert np.array_equal(back.snrs, pgram.snrs)
    assert np.array_equal(back.periods, pgram.periods)
    assert np.array_equal(back.widths, pgram.widths)
    assert back.metadata == pgram.metadata


def test_ffa_search_no_downsampling(oracle):
    """tests/test_ffa_search_pgram.py:76-94: period_min = bins_min * tsamp
    (first downsampling factor exactly 1)."""
    np.random.seed(1)
    tsamp = 1e-3
    ts = rt.TimeSeries.generate(200.0, tsamp, 1.0, amplitude=20.0)
    kw = dict(period_min=800 * tsamp, period_max=1200 * tsamp, bins_min=800, bins_max=1200)
    tsdr, pgram = rt.ffa_search(ts, **kw)
    xo, wo, Po, Fo, So = oracle.ffa_search(ts.data, tsamp, **kw)
    assert np.array_equal(pgram.periods, Po) and np.array_equal(pgram.foldbins, Fo)
    np.testing.assert_allclose(pgram.snrs, So, rtol=1e-5, atol=1e-4)
