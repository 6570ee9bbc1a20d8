"""GPU parity of the search path (dereddening, normalisation, periodogram,
ffa_search) against the CPU oracle on the same seeded inputs.

Tolerances: trial periods and fold bins bit-exact; S/N within 1e-5 relative
and 1e-4 absolute of the oracle; dereddened / normalised series within a few
float32 ulps (the reference rounds float64 numpy results to float32, the GPU
rounds float-float results once)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SNR_RTOL, SNR_ATOL = 1e-5, 1e-4


def _pulsar(n, tsamp, period, ducy, amplitude, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) * tsamp
    phase = (t / period) % 1.0
    sig = np.exp(-0.5 * ((phase - 0.5) / (ducy / 2.355)) ** 2)
    sig *= amplitude / np.sqrt((sig ** 2).sum())
    return (sig + rng.standard_normal(n)).astype(np.float32)


PGRAM_CASES = [
    # N, tsamp, pmin, pmax, bmin, bmax
    (2 ** 17, 256e-6, 0.5, 2.0, 240, 260),
    (2 ** 17, 256e-6, 0.5, 2.0, 480, 520),
    (200_000, 1e-3, 0.8, 1.2, 240, 260),
    (200_000, 1e-3, 0.8, 1.2, 800, 1200),   # period_min = bins_min * tsamp: first step f == 1
    (2 ** 20, 64e-6, 0.1, 1.0, 480, 520),   # multi-pass transforms (up to ~1300 rows)
    (2 ** 21, 64e-6, 6.0, 12.0, 32, 36),    # downsampling factors 2930-5860 (direct above 4000)
    (2 ** 20, 64e-6, 0.1, 10.0, 240, 241),  # 555 downsampling steps (three packed launches)
]


@pytest.mark.parametrize("case", PGRAM_CASES)
def test_periodogram_parity(oracle, case):
    from riptide_amd import libcpp
    N, tsamp, pmin, pmax, bmin, bmax = case
    x = _pulsar(N, tsamp, 0.3711 * (pmin + pmax), 0.03, 15.0, seed=N % 1000)
    x = oracle.normalise(x)
    w = oracle.generate_width_trials(bmin)
    P, F, S = libcpp.periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    assert P.shape == Po.shape and S.shape == So.shape
    assert np.array_equal(P, Po)
    assert np.array_equal(F, Fo)
    np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)


@pytest.mark.parametrize("case", [(2 ** 18, 128e-6, 0.2, 2.0, 240, 260),
                                  (2 ** 21, 64e-6, 6.0, 12.0, 32, 36),
                                  (2 ** 20, 64e-6, 0.1, 10.0, 240, 241)])
def test_plain_schedule_matches(oracle, case):
    """The pass-by-pass schedule and the batch of one give bit-identical
    results (the same task code over the same downsampled series), which
    equal the oracle's: periods and fold bins exactly, S/N within the
    tolerance.  The third case has more downsampling steps than one packed
    launch holds."""
    from riptide_amd import _lib, libcpp
    N, tsamp, pmin, pmax, bmin, bmax = case
    x = oracle.normalise(_pulsar(N, tsamp, 0.77, 0.02, 20.0, seed=5))
    w = oracle.generate_width_trials(bmin)
    d = _lib.to_device(x)
    a = [_lib.to_numpy(u) for u in
         libcpp.periodogram_device(d, N, tsamp, w, pmin, pmax, bmin, bmax, plain=False)]
    b = [_lib.to_numpy(u) for u in
         libcpp.periodogram_device(d, N, tsamp, w, pmin, pmax, bmin, bmax, plain=True)]
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    assert np.array_equal(a[0], Po) and np.array_equal(a[1], Fo)
    np.testing.assert_allclose(a[2], So, rtol=SNR_RTOL, atol=SNR_ATOL)


def test_periodogram_rows_wider_than_the_tile(oracle):
    """Rows of 3000-3100 bins (47-49 columns per lane) do not fit the staging
    of workgroup shapes 0 and 1 (40 columns per lane): the batch of one
    tiles them with shape 2 (72 columns per lane, rows up to 4095 bins) and
    equals the oracle, also with shape 0 asked for; the pass-by-pass
    schedule, which keeps the shape asked for, raises the library's error."""
    from riptide_amd import _lib, libcpp, tune
    N, tsamp, pmin, pmax, bmin, bmax = 2 ** 18, 256e-6, 0.8, 3.0, 3000, 3100
    x = oracle.normalise(_pulsar(N, tsamp, 1.3, 0.03, 15.0, seed=3000))
    w = oracle.generate_width_trials(bmin)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    for shape in (None, 0):
        tune.set_shape(shape)
        try:
            P, F, S = libcpp.periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
        finally:
            tune.set_shape(None)
        assert np.array_equal(P, Po) and np.array_equal(F, Fo)
        np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)
    tune.set_shape(0)
    try:
        with pytest.raises(ValueError, match="bins too large for the FFA tiling"):
            libcpp.periodogram_device(_lib.to_device(x), N, tsamp, w, pmin, pmax, bmin, bmax,
                                      plain=True)
    finally:
        tune.set_shape(None)


def _ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.abs(a.astype(np.float64) - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


@pytest.mark.parametrize("n,tsamp,width,minpts", [
    (2 ** 18, 256e-6, 4.0, 101),   # scrunch factor 154
    (100_000, 1e-3, 0.5, 51),      # scrunch factor 9
    (3000, 1e-3, 0.051, 101),      # no scrunching: exact running median (51 samples)
])
def test_deredden_parity(oracle, n, tsamp, width, minpts):
    from riptide_amd.time_series import deredden_array
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 3 * np.sin(np.arange(n) / 5000.0)).astype(np.float32)
    got = deredden_array(x, int(round(width / tsamp)), minpts)
    want = oracle.deredden(x, tsamp, width, minpts)
    # rounding of a float64 (reference) vs float-float (GPU) difference
    scale = np.abs(x).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=2 * np.spacing(np.float32(scale)))
    assert np.mean(got == want) > 0.99


@pytest.mark.parametrize("sf,extra", [(4, 0), (5, 3), (9, 8), (154, 77), (618, 1)])
def test_deredden_ramp_exact(oracle, sf, extra):
    """A ramp dereddens to exactly zero between the first and last block
    centre (i*sf + (sf-1)/2) and to the ramp minus the end blocks' medians
    outside, also with a trailing partial block (N % sf != 0): every value
    is an integer or half-integer, so the GPU result is exact."""
    from riptide_amd.time_series import deredden_array
    minpts, nl = 11, 40
    n = nl * sf + extra
    x = (2.0 * np.arange(n) - 300.0).astype(np.float32)
    width = sf * minpts + minpts // 2
    got = deredden_array(x, width, minpts)
    want = oracle.deredden(x, 1.0, float(width), minpts)
    assert np.array_equal(got, want)
    c0, c1 = (sf - 1) // 2, (nl - 1) * sf + (sf - 1) // 2
    assert np.all(got[c0 + 1:c1] == 0.0)


@pytest.mark.parametrize("n,sf,minpts", [(5003, 7, 11), (20011, 154, 51), (1_000_003, 618, 101)])
def test_deredden_trailing_block_parity(oracle, n, sf, minpts):
    """Red noise with N % sf != 0: the trailing partial block is dropped
    from the block means and the values past the last centre take the last
    block's median."""
    from riptide_amd.time_series import deredden_array
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 5 * np.sin(np.arange(n) / (3.0 * sf))).astype(np.float32)
    width = sf * minpts + 1
    got = deredden_array(x, width, minpts)
    want = oracle.deredden(x, 1.0, float(width), minpts)
    scale = np.abs(x).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=2 * np.spacing(np.float32(scale)))
    assert np.mean(got == want) > 0.99


def test_normalise_parity(oracle):
    from riptide_amd.time_series import normalise_array
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(1_000_003) * 3 + 42).astype(np.float32)
    got = normalise_array(x)
    want = oracle.normalise(x)
    # the mean's last float64 bits differ (pairwise float64 vs float-float),
    # which moves values near zero by ~1e-13 absolute: compare on the O(1) scale
    np.testing.assert_allclose(got, want, rtol=2.4e-7, atol=2.4e-7)
    assert np.mean(got == want) > 0.99


def test_ffa_search_parity(oracle):
    import riptide_amd as rt
    N, tsamp = 2 ** 18, 256e-6
    x = _pulsar(N, tsamp, 1.3711, 0.02, 20.0, seed=1)
    ts = rt.TimeSeries(x, tsamp)
    tsdr, pg = rt.ffa_search(ts, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                             rmed_width=4.0)
    xo, wo, Po, Fo, So = oracle.ffa_search(x, tsamp, period_min=0.5, period_max=2.0,
                                           bins_min=240, bins_max=260, rmed_width=4.0)
    assert np.array_equal(pg.widths, wo)
    assert np.array_equal(pg.periods, Po)
    assert np.array_equal(pg.foldbins, Fo)
    np.testing.assert_allclose(tsdr.data, xo, rtol=0, atol=4e-6)
    np.testing.assert_allclose(pg.snrs, So, rtol=SNR_RTOL, atol=SNR_ATOL)
    assert pg.metadata == ts.metadata == tsdr.metadata
    assert pg.tobs == ts.length


def test_rseek_end_to_end_pin():
    """tests/test_rseek.py:33-62 of the reference: seed-0 fake pulsar, P = 1 s,
    tobs 128 s, tsamp 256 us, amplitude 20, duty cycle 2%, searched over
    0.5-2 s with 480-520 bins: best S/N 18.5 +/- 0.15 at width 13 within
    0.1/tobs of 1 Hz."""
    import riptide_amd as rt
    np.random.seed(0)
    ts = rt.TimeSeries.generate(128.0, 256e-6, 1.0, amplitude=20.0, ducy=0.02, stdnoise=1.0)
    _, pg = rt.ffa_search(ts, period_min=0.5, period_max=2.0, bins_min=480, bins_max=520)
    assert pg.snrs.shape == (95033, 12)
    sel = np.abs(pg.freqs - 1.0) < 0.1 / 128.0
    best = pg.snrs[sel].max(axis=0)
    assert abs(best.max() - 18.5) < 0.15
    assert pg.widths[best.argmax()] == 13


def test_ffa_search_returns_same_object():
    import riptide_amd as rt
    np.random.seed(1)
    ts = rt.TimeSeries.generate(20.0, 1e-3, 1.0, amplitude=20.0)
    tsdr, pg = rt.ffa_search(ts, period_min=0.8, period_max=1.2, bins_min=240, bins_max=260,
                             already_normalised=True, deredden=False)
    assert id(tsdr) == id(ts)
    assert all(np.maximum.accumulate(pg.periods) == pg.periods)


# Tiles of many short rows (every lane of the S/N phase gets a (chunk, row)
# item from several chunks per row) and width lists other than the default
# trials (the S/N phase's generic window path), under every workgroup shape.
SHAPE_CASES = [
    # N, tsamp, pmin, pmax, bmin, bmax, widths (None: default trials)
    (2 ** 16, 1e-3, 0.02, 0.2, 16, 20, None),
    (2 ** 16, 1e-3, 0.05, 0.5, 40, 44, None),
    (2 ** 17, 256e-6, 0.5, 2.0, 240, 260, [1, 3, 5, 8, 21, 40]),
    (2 ** 17, 256e-6, 0.5, 2.0, 480, 520, [2, 7]),
    # widths out of order, and windows up to nearly a whole row (most window
    # starts then wrap around the row)
    (2 ** 17, 256e-6, 0.5, 2.0, 240, 260, [13, 1, 40, 3, 7]),
    (2 ** 16, 1e-3, 0.05, 0.5, 40, 44, [5, 39, 1, 17]),
]


def _periodogram(x, tsamp, w, pmin, pmax, bmin, bmax, plain):
    """The drop-in periodogram (a batch of one series through the batch
    kernel) or, with plain, the per-step schedule of the task kernel."""
    from riptide_amd import _lib, libcpp
    d = _lib.to_device(np.ascontiguousarray(x, np.float32))
    P, F, S = libcpp.periodogram_device(d, x.size, tsamp, np.asarray(w, np.uint64), pmin, pmax,
                                        bmin, bmax, plain=plain)
    return _lib.to_numpy(P), _lib.to_numpy(F).view(np.uint32), _lib.to_numpy(S)


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("case", SHAPE_CASES)
def test_periodogram_parity_shapes(oracle, case, shape, plain):
    from riptide_amd import tune
    N, tsamp, pmin, pmax, bmin, bmax, widths = case
    x = oracle.normalise(_pulsar(N, tsamp, 0.37 * (pmin + pmax), 0.05, 15.0, seed=bmin + shape))
    w = oracle.generate_width_trials(bmin) if widths is None else np.asarray(widths)
    tune.set_shape(shape)
    try:
        P, F, S = _periodogram(x, tsamp, w, pmin, pmax, bmin, bmax, plain)
    finally:
        tune.set_shape(None)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    assert np.array_equal(P, Po) and np.array_equal(F, Fo)
    np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)


# The S/N phase's compile-time path for the default width trials holds one
# instantiation per width count (7-16): the sliding block serves the first
# min(N, 8) widths and the rest load their own values.  These bins ranges give
# N = 7, 8, 9, 11, 14, 15 and 16 default widths (240, 480 and 800 give 10, 12
# and 13 above); every workgroup shape.
DEFAULT_WIDTH_CASES = [
    # bins_min, samples -> number of default widths
    (70, 2 ** 17),     # 7: fewer widths than the block serves
    (100, 2 ** 17),    # 8: exactly the block, no other widths
    (150, 2 ** 17),    # 9
    (320, 2 ** 17),    # 11
    (1100, 2 ** 18),   # 14
    (1600, 2 ** 18),   # 15
    (2400, 2 ** 18),   # 16: all of kDefaultWidths
]


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("bmin,N", DEFAULT_WIDTH_CASES)
def test_periodogram_default_width_counts(oracle, bmin, N, shape, plain):
    from riptide_amd import tune
    tsamp = 1e-3
    pmin = bmin * tsamp * 1.05
    pmax = 1.6 * pmin
    bmax = bmin + max(4, bmin // 20)
    w = oracle.generate_width_trials(bmin)
    assert w.size == {70: 7, 100: 8, 150: 9, 320: 11, 1100: 14, 1600: 15, 2400: 16}[bmin]
    x = oracle.normalise(_pulsar(N, tsamp, 1.3 * pmin, 0.04, 15.0, seed=bmin + 10 * shape))
    tune.set_shape(shape)
    try:
        P, F, S = _periodogram(x, tsamp, w, pmin, pmax, bmin, bmax, plain)
    finally:
        tune.set_shape(None)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    assert np.array_equal(P, Po) and np.array_equal(F, Fo)
    np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)


def test_ffa_search_results_stay_on_device(oracle):
    """ffa_search leaves the searched series and the S/N in HBM: no host copy
    until they are read, and the copies then equal the oracle's results."""
    import riptide_amd as rt
    N, tsamp = 2 ** 17, 256e-6
    x = _pulsar(N, tsamp, 0.9, 0.03, 15.0, seed=8)
    ts = rt.TimeSeries(x, tsamp)
    tsdr, pg = rt.ffa_search(ts, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                             rmed_width=4.0)
    assert pg.snrs_device is not None and not pg.snrs_on_host
    assert tsdr.data_device is not None and tsdr.nsamp == N
    assert tsdr.metadata["tobs"] == ts.length
    xo, _, Po, Fo, So = oracle.ffa_search(x, tsamp, period_min=0.5, period_max=2.0,
                                          bins_min=240, bins_max=260, rmed_width=4.0)
    peaks, _ = rt.find_peaks(pg)  # reads the device copy
    assert not pg.snrs_on_host
    np.testing.assert_allclose(pg.snrs, So, rtol=SNR_RTOL, atol=SNR_ATOL)
    assert pg.snrs_on_host and tsdr.data_device is not None
    np.testing.assert_allclose(tsdr.data, xo, rtol=0, atol=4e-6)
    assert tsdr.data_device is None
    # a device-backed series searched again: no host round trip needed
    _, pg2 = rt.ffa_search(tsdr, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                           deredden=False, already_normalised=True)
    np.testing.assert_array_equal(pg2.snrs, pg.snrs)


def test_dropin_periodogram_kept_batch(oracle):
    """The drop-in periodogram keeps the batches of its last four searches:
    a second series with the same search (a kept batch, rebound), searches
    that alternate (each kept), more searches than are kept (the least
    recently used dropped) and an evicted search again (a new batch) all
    equal the oracle."""
    N, tsamp = 2 ** 17, 256e-6
    a = (0.5, 2.0, 240, 260)
    b = (0.4, 1.6, 480, 520)
    c = (0.5, 1.9, 240, 260)
    d = (0.6, 2.0, 250, 270)
    e = (0.5, 1.8, 480, 520)
    x1 = oracle.normalise(_pulsar(N, tsamp, 0.83, 0.03, 15.0, seed=21))
    x2 = oracle.normalise(_pulsar(N, tsamp, 1.27, 0.03, 15.0, seed=22))
    calls = ((x1, a), (x2, a), (x1, b), (x2, a), (x1, b), (x2, c), (x1, d), (x2, e), (x1, a),
             (x2, e), (x1, b))
    for x, (pmin, pmax, bmin, bmax) in calls:
        w = oracle.generate_width_trials(bmin)
        P, F, S = _periodogram(x, tsamp, w, pmin, pmax, bmin, bmax, False)
        Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
        assert np.array_equal(P, Po) and np.array_equal(F, Fo)
        np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)


def test_release_scratch(oracle):
    """release_scratch gives back the kept batches of the drop-in periodogram
    (device memory rises by about its size) and later searches still equal
    the oracle."""
    import torch
    import riptide_amd as rt
    N, tsamp, pmin, pmax, bmin, bmax = 2 ** 18, 256e-6, 0.5, 2.0, 480, 520
    x = oracle.normalise(_pulsar(N, tsamp, 0.91, 0.03, 15.0, seed=31))
    w = oracle.generate_width_trials(bmin)
    Po, Fo, So = oracle.oracle().periodogram(x, tsamp, w, pmin, pmax, bmin, bmax)
    for _ in range(2):
        P, F, S = _periodogram(x, tsamp, w, pmin, pmax, bmin, bmax, False)
        assert np.array_equal(P, Po) and np.array_equal(F, Fo)
        np.testing.assert_allclose(S, So, rtol=SNR_RTOL, atol=SNR_ATOL)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rt.release_scratch()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 > free0
