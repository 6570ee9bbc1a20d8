"""Candidate folding on the GPU (TimeSeries.fold, downsample_vertical)
against the oracle restatement: the reference's shapes, the float32 scaling
and sums, downsampling within its float32 tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("period,bins,subints", [
    (1.3711, 128, None), (1.3711, 128, 1), (1.3711, 250, 16), (0.5, 64, 7), (2.0, 100, 3),
    (0.0123, 8, None), (1.3711, 128, 191),
])
def test_fold_matches_oracle(period, bins, subints):
    import riptide_amd as rt
    from oracle import folding as OF
    rng = np.random.default_rng(int(period * 1000) + bins)
    n, tsamp = 2 ** 18, 1e-3
    x = rng.standard_normal(n).astype(np.float32)
    ts = rt.TimeSeries.from_numpy_array(x, tsamp)
    got = ts.fold(period, bins, subints=subints)
    want = OF.fold(x, tsamp, period, bins, subints=subints)
    assert got.dtype == np.float32 and got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5)


def test_downsample_vertical_matches_oracle():
    from oracle import folding as OF
    from riptide_amd.folding import downsample_vertical
    X = np.random.default_rng(1).standard_normal((301, 77)).astype(np.float32)
    for f in (1.5, 2.0, 7.3, 300.5):
        got = downsample_vertical(X, f)
        want = OF.downsample_vertical(X, f)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5)
