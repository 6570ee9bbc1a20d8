"""Host parts of the device peak detection against numpy itself and known
answers: the percentile interpolation the device path rebuilds from order
statistics equals np.percentile; cluster1d on hand-made cases; the oracle
finds planted peaks and falls back to the constant threshold below minseg."""
import numpy as np
import pytest

from riptide_amd.clustering import cluster1d
from riptide_amd.peak_detection import lerp, percentile_ranks, _smin_lo


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 7, 8, 100, 1237, 2500, 4097])
def test_percentile_from_order_statistics_equals_numpy(p):
    rng = np.random.default_rng(p)
    for trial in range(5):
        x = rng.standard_normal(p).astype(np.float32).astype(np.float64)
        if trial == 1:
            x = np.round(x, 1)  # ties
        srt = np.sort(x)
        want = np.percentile(x, (25, 50, 75))
        got = [float(lerp(srt[lo], srt[hi], g)) for lo, hi, g in percentile_ranks(p)]
        assert np.array_equal(got, want), (p, trial, got, want)


def test_cluster1d_known_answers():
    x = np.array([5.0, 1.0, 1.05, 9.0, 5.02, 1.1, 20.0])
    cl = cluster1d(x, 0.06)
    assert [sorted(c.tolist()) for c in cl] == [[1, 2, 5], [0, 4], [3], [6]]
    assert cluster1d(np.array([]), 1.0) == []
    one = cluster1d(np.array([3.0, 3.1, 3.15]), 0.2)
    assert len(one) == 1 and sorted(one[0].tolist()) == [0, 1, 2]
    srt = cluster1d(np.array([1.0, 2.0, 4.0]), 1.5, already_sorted=True)
    assert [c.tolist() for c in srt] == [[0, 1], [2]]


@pytest.mark.parametrize("smin", [6.0, 6.1, 5.999999, 7.25])
def test_smin_float32_bound(smin):
    lo = _smin_lo(smin)
    assert lo <= smin and np.float32(lo) == lo
    nxt = float(np.nextafter(np.float32(lo), np.float32(np.inf)))
    assert nxt > smin or lo == smin


def _planted(L=20000, nw=3, seed=0):
    rng = np.random.default_rng(seed)
    periods = np.sort(rng.uniform(0.5, 2.0, L))[::-1].copy()  # increasing frequency
    snrs = (rng.standard_normal((L, nw)) + 3.0).astype(np.float32)
    peaks_at = [1234, 9000, 15555]
    for ip in peaks_at:
        snrs[ip, :] += 25.0
    foldbins = np.full(L, 250, np.uint32)
    return periods, foldbins, np.array([1, 2, 3]), snrs, peaks_at


def test_oracle_finds_planted_peaks():
    from oracle import peaks as OP
    periods, foldbins, widths, snrs, at = _planted()
    tobs = 300.0
    pk, polycos = OP.find_peaks(periods, foldbins, widths, snrs, tobs)
    found = sorted({p[5] for p in pk})
    assert found == sorted(at)
    assert all(len(polycos[iw]) == 3 for iw in range(3))
    # fewer segments than minseg: the constant threshold smin, as a list
    pk2, polycos2 = OP.find_peaks(periods, foldbins, widths, snrs, tobs, smin=12.0,
                                  minseg=10 ** 6)
    assert polycos2[0] == [12.0]
    assert sorted({p[5] for p in pk2}) == sorted(at)


def test_fold_argument_errors_match_reference():
    """folding.py:52-67 checks, raised before any device work."""
    import riptide_amd as rt
    from riptide_amd.folding import downsample_vertical
    ts = rt.TimeSeries.from_numpy_array(np.zeros(1000, np.float32), 1e-3)
    with pytest.raises(ValueError, match="Period exceeds data length"):
        ts.fold(2.0, 10)
    with pytest.raises(ValueError, match="Bin width is shorter than sampling time"):
        ts.fold(0.05, 100)
    with pytest.raises(ValueError, match="subints must be >= 1 or None"):
        ts.fold(0.1, 10, subints=0)
    with pytest.raises(ValueError, match="exceeds the number of signal periods"):
        ts.fold(0.3, 10, subints=4)
    with pytest.raises(ValueError, match="factor must be > 1"):
        downsample_vertical(np.zeros((5, 3)), 1.0)
    with pytest.raises(ValueError, match="strictly smaller"):
        downsample_vertical(np.zeros((5, 3)), 5.0)
