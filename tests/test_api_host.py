"""The reference's own API tests restated against riptide_amd, for the parts
that need no GPU: the ffafreq / ffaprd closed form and its m = 1 case, and
every argument error of ffa1, ffa2, boxcar_snr, snr1, snr2, downsample,
running_median and fast_running_median, which are raised on the host before
any device work (tests/test_ffa_base_functions.py:39-118,
tests/test_snr.py:7-22, tests/test_running_median.py:20-56,
tests/test_time_series.py:160-164 of the reference; messages from
python_bindings.cpp:33-41, snr.hpp:17,30, running_median.hpp:104,107,
downsample.hpp:15, libffa.py:119-159 and running_medians.py:73-74)."""
import numpy as np
import pytest
from pytest import raises

import riptide_amd as rt
from riptide_amd import libcpp


# ---------------------------------------------------------------------------
# ffafreq / ffaprd (tests/test_ffa_base_functions.py:84-118)
# ---------------------------------------------------------------------------

def test_ffafreq_closed_form():
    """P_s = p^2 / (p - s/(m-1)) * dt for the rows s of an m x p transform
    (arXiv:2004.03701); ffaprd is its reciprocal."""
    m, p = 42, 127
    N = m * p
    dt = np.pi / 1000.0
    s = np.arange(m, dtype=float)
    true_periods = p ** 2 / (p - s / (m - 1.0)) * dt
    freqs = rt.ffafreq(N, p, dt=dt)
    periods = rt.ffaprd(N, p, dt=dt)
    np.testing.assert_allclose(freqs, 1.0 / true_periods, rtol=1e-14)
    np.testing.assert_allclose(periods, true_periods, rtol=1e-14)
    assert freqs.dtype == np.float64 and freqs.shape == (m,)
    # row 0 folds at the base period, row m-1 at p^2/(p-1) samples
    assert np.isclose(periods[0], p * dt) and np.isclose(periods[-1], p ** 2 / (p - 1) * dt)
    # the trailing N % p samples form no row
    assert rt.ffafreq(N + p - 1, p, dt=dt).shape == (m,)


def test_ffafreq_single_period():
    """m = 1: the one row folds at the base period (libffa.py:163-164)."""
    p, dt = 127, np.pi / 1000.0
    f = rt.ffafreq(p, p, dt=dt)
    assert f.shape == (1,) and f[0] == 1.0 / (p * dt)
    assert rt.ffaprd(p + 5, p, dt=dt)[0] == 1.0 / (1.0 / (p * dt))


@pytest.mark.parametrize("args,kw,msg", [
    ((0, 127), {}, "N must be a strictly positive integer"),
    ((np.pi, 127), {}, "N must be a strictly positive integer"),
    ((5334, 1), {}, "p must be an integer > 1"),
    ((5334, np.pi), {}, "p must be an integer > 1"),
    ((5334, 5335), {}, "p must be smaller than"),
    ((5334, 127), {"dt": 0}, "dt must be strictly positive"),
])
def test_ffafreq_errors(args, kw, msg):
    with raises(ValueError, match=msg):
        rt.ffafreq(*args, **kw)
    with raises(ValueError, match=msg):
        rt.ffaprd(*args, **kw)


# ---------------------------------------------------------------------------
# ffa1 / ffa2 argument errors (tests/test_ffa_base_functions.py:59-81)
# ---------------------------------------------------------------------------

def test_ffa2_rejects_1d():
    with raises(ValueError, match="incorrect number of dimensions: 1; expected 2"):
        rt.ffa2(np.zeros(4))


def test_ffa2_rejects_non_contiguous_float32():
    x = np.zeros((8, 16), np.float32)[:, ::2]
    with raises(ValueError, match="Input numpy array must be contiguous in memory"):
        libcpp.ffa2(x)


def test_ffa1_errors():
    with raises(ValueError, match="one-dimensional"):
        rt.ffa1(np.zeros((4, 4)), 4)
    X = np.zeros(10)
    with raises(ValueError, match="smaller than the total number of samples"):
        rt.ffa1(X, X.size + 1)
    with raises(ValueError, match="p must be an integer > 1"):
        rt.ffa1(X, 4.0)
    with raises(ValueError, match="p must be an integer > 1"):
        rt.ffa1(X, 0)
    with raises(ValueError, match="p must be an integer > 1"):
        rt.ffa1(X, True)


# ---------------------------------------------------------------------------
# boxcar_snr / snr1 / snr2 (tests/test_snr.py:7-22)
# ---------------------------------------------------------------------------

def test_boxcar_snr_errors():
    cols = 32
    data = np.zeros(cols, dtype="float32")
    with raises(ValueError, match="trial widths must be all > 0 and < columns"):
        rt.boxcar_snr(data, [0, 1])
    with raises(ValueError, match="trial widths must be all > 0 and < columns"):
        rt.boxcar_snr(data, [1, cols])
    with raises(ValueError, match="stdnoise must be > 0"):
        rt.boxcar_snr(data, [1, 2], stdnoise=-42.0)


def test_snr_error_order():
    """stdnoise is checked before the widths (python_bindings.cpp:131-132),
    contiguity and dimensions before both (python_bindings.cpp:124-129)."""
    x = np.zeros((3, 32), np.float32)
    with raises(ValueError, match="stdnoise"):
        libcpp.snr2(x, [0, 99], stdnoise=0.0)
    with raises(ValueError, match="stdnoise"):
        libcpp.snr1(x[0], [0, 99], stdnoise=float("nan"))
    with raises(ValueError, match="contiguous"):
        libcpp.snr2(np.zeros((3, 64), np.float32)[:, ::2], [1, 2], stdnoise=-1.0)
    with raises(ValueError, match="contiguous"):
        libcpp.snr2(x, np.arange(8, dtype=np.uint64)[::2], stdnoise=-1.0)
    with raises(ValueError, match="incorrect number of dimensions: 2; expected 1"):
        libcpp.snr1(x, [1, 2], stdnoise=-1.0)
    with raises(ValueError, match="incorrect number of dimensions: 1; expected 2"):
        libcpp.snr2(x[0], [1, 2], stdnoise=-1.0)


def test_snr_empty_outputs():
    """No rows or no widths: an empty result of the right shape, after the
    argument checks."""
    assert libcpp.snr2(np.zeros((0, 32), np.float32), [1, 2]).shape == (0, 2)
    assert libcpp.snr2(np.zeros((3, 32), np.float32), []).shape == (3, 0)
    with raises(ValueError, match="trial widths"):
        libcpp.snr2(np.zeros((0, 32), np.float32), [32])


# ---------------------------------------------------------------------------
# running medians (tests/test_running_median.py:20-56)
# ---------------------------------------------------------------------------

def test_running_median_errors():
    data = np.arange(10, dtype="float32")
    with raises(ValueError, match="width must be an odd number"):
        rt.running_median(data, 2)
    with raises(ValueError, match="width must be an odd number"):
        rt.running_median(data, data.size)  # 10: the odd check comes first
    with raises(ValueError, match="width must be < size"):
        rt.running_median(data, data.size + 1)
    with raises(ValueError, match="incorrect number of dimensions: 2; expected 1"):
        rt.running_median(np.zeros(shape=(4, 8)), 3)


def test_fast_running_median_min_points_odd():
    with raises(ValueError, match="min_points must be an odd number"):
        rt.fast_running_median(np.zeros(100), 3, min_points=10)


# ---------------------------------------------------------------------------
# downsample (tests/test_time_series.py:160-164, downsample.hpp:11-16)
# ---------------------------------------------------------------------------

def test_downsample_errors():
    x = np.zeros(1000, np.float32)
    for f in (0.55, 1.0, 1000.5, 10000):
        with raises(ValueError, match="Downsampling factor must verify: 1 < f <= size"):
            rt.downsample(x, f)
    ts = rt.TimeSeries(x, 1e-3)
    with raises(ValueError):
        ts.downsample(0.55)
    with raises(ValueError):
        ts.downsample(ts.nsamp * 10)


# ---------------------------------------------------------------------------
# containers (tests/test_time_series.py:63-99, test_ffa_search_pgram.py:30-35)
# ---------------------------------------------------------------------------

def test_time_series_constructors(tmp_path):
    ref = np.arange(16)
    tsamp = 64e-6
    for ts in (rt.TimeSeries.from_numpy_array(ref, tsamp),):
        assert ts.nsamp == 16 and ts.tsamp == tsamp and ts.data.dtype == np.float32
        assert np.array_equal(ts.data, ref)
    np.save(tmp_path / "x.npy", ref)
    ts = rt.TimeSeries.from_npy_file(str(tmp_path / "x.npy"), tsamp)
    assert ts.nsamp == 16 and np.array_equal(ts.data, ref)
    ref.astype(np.float32).tofile(tmp_path / "x.bin")
    ts = rt.TimeSeries.from_binary(str(tmp_path / "x.bin"), tsamp)
    assert ts.nsamp == 16 and np.array_equal(ts.data, ref)
    assert ts.metadata["tobs"] == ts.length == 16 * tsamp


def test_generate_amplitude():
    """Noiseless fake pulsar: L2 norm equals the amplitude (libffa.py:57-59)."""
    ts = rt.TimeSeries.generate(10.0, 0.01, 1.0, amplitude=25.0, stdnoise=0)
    assert ts.length == 10.0 and ts.tsamp == 0.01 and ts.data.dtype == np.float32
    assert np.isclose(np.sum(ts.data.astype(np.float64) ** 2) ** 0.5, 25.0, atol=1e-5)


def test_periodogram_container():
    P = np.array([1.0, 1.5, 2.0])
    pg = rt.Periodogram(np.array([1, 2]), P, np.array([240, 240, 241], np.uint32),
                        np.zeros((3, 2), np.float32), metadata=rt.Metadata({"tobs": 12.5}))
    assert np.all(pg.freqs == 1.0 / pg.periods)
    assert pg.tobs == 12.5
    back = rt.Periodogram.from_dict(pg.to_dict())
    assert np.array_equal(back.snrs, pg.snrs) and back.metadata == pg.metadata
