"""Pins the CPU oracle before it is trusted as the GPU checker.

* against the compiled reference (oracle/_ref, built from the reference's own
  headers) bit for bit, when that library is present;
* against the reference tests' known answers (8x8 FFA table, top-hat S/N,
  naive running median, ffafreq closed form, plan sizes);
* against independent restatements: integer brute force of the FFA from
  per-row shift bookkeeping, float64 boxcar correlation, float64 integration
  of the downsampling step function.
"""
import numpy as np
import pytest


def _ref_or_skip(oracle):
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return oracle.ref()


# ---------------------------------------------------------------------------
# FFA
# ---------------------------------------------------------------------------

# tests/test_ffa_base_functions.py:12-36 of the reference (data, restated)
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


def test_ffa_kat_8x8(oracle):
    o = oracle.oracle()
    for shift in range(8):
        x = np.roll(FFA_IN_88, shift, axis=1)
        assert np.array_equal(o.ffa2(x), np.roll(FFA_OUT_88, shift, axis=1))
    for extra in range(8):
        x = np.hstack([FFA_IN_88, np.zeros((8, extra), np.float32)])
        want = np.hstack([FFA_OUT_88, np.zeros((8, extra), np.float32)])
        assert np.array_equal(o.ffa2(x), want)


def _row_map(k32, s):
    # trunc(fma(k, s, 0.5f)): k*s is exact in float64, so rounding the float64
    # sum once to float32 is the fused multiply-add
    return int(np.float32(np.float64(k32) * s + 0.5))


def _shift_table(m):
    """S[s][r]: left roll applied to input row r in output row s, by following
    the head/tail recursion of transforms.hpp:30-50 on shifts instead of data."""
    if m == 1:
        return [[0]]
    mh = m >> 1
    mt = m - mh
    H = _shift_table(mh)
    T = _shift_table(mt)
    kh = np.float32((np.float32(mh) - np.float32(1.0)) / (np.float32(m) - np.float32(1.0)))
    kt = np.float32((np.float32(mt) - np.float32(1.0)) / (np.float32(m) - np.float32(1.0)))
    out = []
    for s in range(m):
        h = _row_map(kh, s)
        t = _row_map(kt, s)
        out.append(list(H[h]) + [v + (s - t) for v in T[t]])
    return out


@pytest.mark.parametrize("m,p", [(2, 5), (3, 4), (5, 7), (8, 8), (13, 11), (37, 9), (64, 16)])
def test_ffa_integer_brute_force(oracle, m, p):
    rng = np.random.default_rng(m * 100 + p)
    x = rng.integers(-50, 50, size=(m, p)).astype(np.float32)
    S = _shift_table(m)
    want = np.zeros((m, p), np.float64)
    for s in range(m):
        for r in range(m):
            want[s] += np.roll(x[r].astype(np.float64), -S[s][r])
    assert np.array_equal(oracle.oracle().ffa2(x), want.astype(np.float32))


def test_ffa_linearity_and_roll(oracle):
    o = oracle.oracle()
    rng = np.random.default_rng(1)
    a = rng.integers(-9, 9, size=(37, 23)).astype(np.float32)
    b = rng.integers(-9, 9, size=(37, 23)).astype(np.float32)
    assert np.array_equal(o.ffa2(3 * a - 2 * b), 3 * o.ffa2(a) - 2 * o.ffa2(b))
    y = rng.standard_normal((53, 31)).astype(np.float32)
    for k in (1, 7, 30):
        # the same operands are added, only in other columns: exact for floats
        assert np.array_equal(o.ffa2(np.roll(y, k, axis=1)), np.roll(o.ffa2(y), k, axis=1))


def test_row_map_contraction_insensitive():
    """The float32 row map gives the same indices with or without the fused
    multiply-add for every m up to 6000 (both appear in reference builds)."""
    for m in range(2, 6001):
        mh = m >> 1
        for mc in (mh, m - mh):
            k = np.float32((np.float32(mc) - np.float32(1.0)) / (np.float32(m) - np.float32(1.0)))
            s = np.arange(m, dtype=np.float64)
            fused = np.float32(np.float64(k) * s + 0.5).astype(np.int64)
            split = (np.float32(np.float32(k) * s.astype(np.float32)) + np.float32(0.5)).astype(np.int64)
            assert np.array_equal(fused, split), m


@pytest.mark.parametrize("m,p", [(3, 5), (7, 240), (37, 260), (257, 5), (536, 250),
                                 (2684, 260), (8195, 5), (1, 7), (2, 9)])
def test_ffa_matches_compiled_reference(oracle, m, p):
    ref = _ref_or_skip(oracle)
    x = np.random.default_rng(m + p).standard_normal((m, p)).astype(np.float32)
    assert np.array_equal(oracle.oracle().ffa2(x), ref.ffa2(x))


# ---------------------------------------------------------------------------
# Boxcar S/N
# ---------------------------------------------------------------------------

def test_snr_top_hat_closed_form(oracle):
    """tests/test_snr.py:62-78: a width-w top-hat in 64 bins peaks at trial
    width w with S/N = w*h, h = sqrt((n-w)/(n*w))."""
    n = 64
    widths = np.arange(1, n)
    for w in range(1, n):
        x = np.zeros((1, n), np.float32)
        x[0, :w] = 1.0
        snr = oracle.oracle().snr2(x, widths)[0]
        assert snr.argmax() == w - 1
        assert np.isclose(snr.max(), w * np.sqrt((n - w) / (n * w)))


def test_snr_float64_brute_force(oracle):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((20, 97)).astype(np.float32)
    widths = np.array([1, 2, 3, 5, 8, 13, 21, 40])
    got = oracle.oracle().snr2(x, widths, 1.3)
    n = x.shape[1]
    for r in range(x.shape[0]):
        xx = np.concatenate([x[r], x[r]]).astype(np.float64)
        tot = x[r].astype(np.float64).sum()
        for i, w in enumerate(widths):
            h = np.sqrt((n - w) / (n * w))
            b = w / (n - w) * h
            best = max(xx[j:j + w].sum() for j in range(n))
            want = ((h + b) * best - b * tot) / 1.3
            assert abs(got[r, i] - want) < 1e-5 * max(1, abs(want))


def test_snr_rotation_invariance(oracle):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, 32)).astype(np.float32)
    w = [1, 2, 5, 11, 18, 31]
    base = oracle.oracle().snr2(x, w)
    for k in range(1, 33):
        np.testing.assert_allclose(oracle.oracle().snr2(np.roll(x, k, axis=1), w), base,
                                   rtol=1e-5, atol=1e-5)


def test_snr_matches_compiled_reference(oracle):
    ref = _ref_or_skip(oracle)
    x = np.random.default_rng(9).standard_normal((300, 260)).astype(np.float32) * 7
    w = oracle.generate_width_trials(240)
    assert np.array_equal(oracle.oracle().snr2(x, w, 2.5), ref.snr2(x, w, 2.5))


# ---------------------------------------------------------------------------
# Downsampling
# ---------------------------------------------------------------------------

def test_downsample_integer_factor_is_block_sum(oracle):
    x = np.random.default_rng(5).integers(-100, 100, 1000).astype(np.float32)
    for f in (2, 3, 10, 7):
        n = x.size // f
        assert np.array_equal(oracle.oracle().downsample(x, float(f)),
                              x[: n * f].reshape(n, f).sum(axis=1))


def test_downsample_fractional_integral(oracle):
    """out[k] = integral over [kf, (k+1)f) of the step function x(t) = x[floor t]
    (cpp/README.md:3-12), evaluated in float64."""
    x = np.random.default_rng(6).standard_normal(5000).astype(np.float32)
    for f in (1.5, 2.25, 6.51, 13.777):
        got = oracle.oracle().downsample(x, f)
        c = np.concatenate([[0.0], np.cumsum(x.astype(np.float64))])

        def F(t):  # integral of the step function from 0 to t
            i = np.minimum(np.floor(t).astype(np.int64), x.size - 1)
            return c[i] + (t - i) * x[i]

        k = np.arange(got.size)
        want = F(np.minimum((k + 1) * f, x.size)) - F(k * f)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.sqrt(f))


@pytest.mark.parametrize("f", [1.5, 2.0, 6.510416666666667, 13.777, 600.3])
def test_downsample_matches_compiled_reference(oracle, f):
    ref = _ref_or_skip(oracle)
    x = np.random.default_rng(8).standard_normal(100003).astype(np.float32)
    assert np.array_equal(oracle.oracle().downsample(x, f), ref.downsample(x, f))


# ---------------------------------------------------------------------------
# Running median
# ---------------------------------------------------------------------------

def _naive_rmed(x, w):
    # tests/test_running_median.py:7-17 of the reference
    h = w // 2
    p = np.pad(x, (h, h), mode="edge")
    return np.asarray([np.median(p[i:i + w]) for i in range(x.size)], dtype=np.float32)


@pytest.mark.parametrize("w", [1, 3, 5, 7, 11, 25, 37])
def test_running_median_naive(oracle, w):
    x = np.random.default_rng(w).standard_normal(100).astype(np.float32)
    assert np.array_equal(oracle.oracle().running_median(x, w), _naive_rmed(x, w))


def test_running_median_errors(oracle):
    x = np.arange(10, dtype=np.float32)
    with pytest.raises(ValueError, match="odd"):
        oracle.oracle().running_median(x, 2)
    with pytest.raises(ValueError, match="< size"):
        oracle.oracle().running_median(x, 11)


def test_running_median_matches_compiled_reference(oracle):
    ref = _ref_or_skip(oracle)
    x = np.random.default_rng(3).standard_normal(3000).astype(np.float32)
    for w in (1, 3, 101, 555):
        assert np.array_equal(oracle.oracle().running_median(x, w), ref.running_median(x, w))


# ---------------------------------------------------------------------------
# Search plan and periodogram
# ---------------------------------------------------------------------------

# (N, tsamp, pmin, pmax, bmin, bmax) -> number of trial periods
PLAN_LENGTHS = [
    ((2 ** 20, 256e-6, 0.5, 2.0, 240, 260), 100_193),
    ((500_000, 256e-6, 0.5, 2.0, 480, 520), 95_033),
    ((200_000, 1e-3, 0.8, 1.2, 240, 260), 20_712),
    ((2 ** 22, 64e-6, 0.1, 10.0, 240, 260), 662_386),
    ((2 ** 22, 64e-6, 0.1, 10.0, 480, 520), 1_324_675),
    ((2 ** 23, 64e-6, 0.1, 10.0, 480, 520), 2_652_094),
]


@pytest.mark.parametrize("args,length", PLAN_LENGTHS)
def test_plan_lengths(oracle, args, length):
    assert oracle.oracle().periodogram_length(*args) == length
    if oracle.ref_available():
        assert oracle.ref().periodogram_length(*args) == length


def test_period_grid_closed_form(oracle):
    """Each transform's periods follow ffaprd's p^2/(p - s/(m-1)) * tau."""
    x = np.random.default_rng(0).standard_normal(2 ** 16).astype(np.float32)
    P, F, S = oracle.oracle().periodogram(x, 256e-6, [1, 2, 3], 0.5, 2.0, 240, 260)
    assert np.all(np.diff(P) >= 0)
    # first transform: f = ds_ini, bins = 240, rows = n // 240
    f = 0.5 / (256e-6 * 240)
    n = int(np.floor(x.size / f))
    m = n // 240
    run = int(np.argmax(F != 240))  # the first transform is the leading run of 240s
    first = P[:run]
    s = np.arange(first.size)
    np.testing.assert_allclose(first, 240 ** 2 / (240 - s / (m - 1.0)) * f * 256e-6, rtol=1e-14)


def test_periodogram_matches_compiled_reference(oracle):
    ref = _ref_or_skip(oracle)
    x = np.random.default_rng(1).standard_normal(2 ** 17).astype(np.float32)
    w = oracle.generate_width_trials(480)
    a = oracle.oracle().periodogram(x, 256e-6, w, 0.5, 2.0, 480, 520)
    b = ref.periodogram(x, 256e-6, w, 0.5, 2.0, 480, 520)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_periodogram_errors(oracle):
    o = oracle.oracle()
    x = np.zeros(1000, np.float32)
    for args, msg in [((0.0, 1, 2, 10, 20), "tsamp"), ((1.0, 0, 2, 10, 20), "period_min"),
                      ((1.0, 2, 2, 10, 20), "period_max"), ((1.0, 20, 30, 1, 20), "bins_min"),
                      ((1.0, 20, 30, 10, 9), "bins_max"), ((1.0, 5, 30, 10, 20), "tsamp \\* bins_min")]:
        with pytest.raises(ValueError, match=msg):
            o.periodogram_length(x.size, *args)


def test_width_trials():
    from oracle.oracle import generate_width_trials
    assert generate_width_trials(240).tolist() == [1, 2, 3, 4, 6, 9, 13, 19, 28, 42]
    assert generate_width_trials(480).tolist() == [1, 2, 3, 4, 6, 9, 13, 19, 28, 42, 63, 94]


def test_rseek_pin_oracle(oracle):
    """The end-to-end pin of tests/test_rseek.py:59-62 holds for the oracle:
    seed-0 fake pulsar gives best S/N 18.5 +/- 0.15 at width 13."""
    np.random.seed(0)
    nsamp = int(round(128.0 / 256e-6))
    P = 1.0 / 256e-6
    kappa = np.log(2.0) / (2.0 * np.sin(np.pi * 0.02 / 2.0) ** 2)
    ph = (np.arange(nsamp, dtype=float) / P - 0.5) * (2 * np.pi)
    sig = np.exp(kappa * (np.cos(ph) - 1.0))
    sig *= 20.0 * (sig ** 2).sum() ** -0.5
    data = (sig + np.random.normal(size=nsamp)).astype(np.float32)
    _, w, periods, _, snrs = oracle.ffa_search(data, 256e-6, period_min=0.5, period_max=2.0,
                                               bins_min=480, bins_max=520)
    sel = np.abs(1.0 / periods - 1.0) < 0.1 / 128.0
    best = snrs[sel].max(axis=0)
    assert abs(best.max() - 18.5) < 0.15 and w[best.argmax()] == 13


def test_deredden_constant_is_zero(oracle):
    x = np.full(10_000, 42.42, np.float32)
    assert np.allclose(oracle.deredden(x, 1e-3, 0.5, 51), 0.0, atol=1e-6)


# ---------------------------------------------------------------------------
# Dereddening with scrunching (running_medians.py:40-85, time_series.py:95-124)
# ---------------------------------------------------------------------------

def numpy_pairwise_sum_f32(a):
    """numpy's pairwise float32 summation (pairwise_sum in numpy's
    umath/loops_utils.h.src, the inner loop of add.reduce over a contiguous
    axis, started from the additive identity): blocks of at most 128 values
    summed with eight interleaved accumulators combined as
    ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the rest one by one; longer
    runs split at a multiple of 8 near the middle."""
    n = a.size
    f = np.float32
    if n < 8:
        r = f(0.0)
        for v in a:
            r = f(r + v)
        return r
    if n <= 128:
        r = [f(v) for v in a[:8]]
        i = 8
        while i < n - n % 8:
            for j in range(8):
                r[j] = f(r[j] + a[i + j])
            i += 8
        res = f(f(f(r[0] + r[1]) + f(r[2] + r[3])) + f(f(r[4] + r[5]) + f(r[6] + r[7])))
        while i < n:
            res = f(res + a[i])
            i += 1
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return f(numpy_pairwise_sum_f32(a[:n2]) + numpy_pairwise_sum_f32(a[n2:]))


@pytest.mark.parametrize("sf", [2, 3, 8, 9, 100, 128, 129, 154, 309, 618])
def test_scrunch_is_numpy_pairwise_mean(oracle, sf):
    """The block means are numpy's float32 pairwise sums divided by the
    float32 factor, bit for bit, including the trailing partial block being
    dropped (N % sf != 0)."""
    rng = np.random.default_rng(sf)
    x = (rng.standard_normal(sf * 11 + sf // 2 + 1) * 100 + 7).astype(np.float32)
    got = oracle.scrunch(x, sf)
    assert got.dtype == np.float32 and got.size == x.size // sf
    want = np.array([numpy_pairwise_sum_f32(x[j * sf:(j + 1) * sf]) / np.float32(sf)
                     for j in range(x.size // sf)], dtype=np.float32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("sf,extra", [(4, 0), (5, 3), (9, 8), (154, 77), (618, 1)])
def test_fast_running_median_reproduces_a_ramp(oracle, sf, extra):
    """On a ramp x[i] = a*i + b the block means are the ramp at the block
    centres i*sf + (sf-1)/2, the running median of a monotonic sequence is the
    sequence itself (edge padding included), and linear interpolation over the
    centres reproduces the ramp between the first and last centre and clamps
    to the end values outside.  A centre off by half a sample (e.g. i*sf +
    sf/2) moves every interior value by a/2 and fails this test; the values
    are integers and half-integers, so every step is exact."""
    minpts = 11
    nl = 40
    n = nl * sf + extra  # the trailing `extra` samples form no block
    a, b = 2.0, -300.0
    x = (a * np.arange(n) + b).astype(np.float32)
    width = sf * minpts + minpts // 2  # int(width / minpts) == sf
    got = oracle.fast_running_median(x, width, minpts)
    assert got.dtype == np.float64 and got.size == n
    c0, c1 = 0.5 * (sf - 1), (nl - 1) * sf + 0.5 * (sf - 1)
    i = np.arange(n, dtype=np.float64)
    want = a * np.clip(i, c0, c1) + b
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * abs(b))
    # the trailing samples past the last centre hold the last block's median
    assert np.all(got[int(np.ceil(c1)):] == got[-1])


@pytest.mark.parametrize("n,sf,minpts", [(5003, 7, 11), (20011, 154, 51), (100_000, 618, 101)])
def test_fast_running_median_interpolation_restated(oracle, n, sf, minpts):
    """Element-by-element check of the interpolation step against an explicit
    float64 linear interpolation over the centres c_j = j*sf + (sf-1)/2
    (index j found by integer arithmetic, not by search), clamped to the end
    values outside [c_0, c_last], on noisy red data with N % sf != 0."""
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 5 * np.sin(np.arange(n) / (3.0 * sf))).astype(np.float32)
    width = sf * minpts + 1
    got = oracle.fast_running_median(x, width, minpts)
    lo = oracle.oracle().running_median(oracle.scrunch(x, sf), minpts).astype(np.float64)
    nl = lo.size
    assert nl == n // sf
    want = np.empty(n)
    for i in range(n):
        # 2*c_j = 2*j*sf + sf - 1: the last centre at or below i
        j = (2 * i - (sf - 1)) // (2 * sf)
        if j < 0:
            want[i] = lo[0]
        elif j >= nl - 1:
            want[i] = lo[-1]
        else:
            t = (i - (j * sf + 0.5 * (sf - 1))) / sf
            want[i] = lo[j] + t * (lo[j + 1] - lo[j])
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14 * np.abs(lo).max())
    # at the centres themselves (integers when sf is odd) the value is the
    # low-resolution median exactly
    if sf % 2:
        centres = np.arange(nl) * sf + (sf - 1) // 2
        assert np.array_equal(got[centres], lo)


def test_deredden_rounds_float64_difference(oracle):
    """time_series.py:121-124 then TimeSeries(...) (time_series.py:41-45):
    float32 data minus the float64 running median, rounded once to float32."""
    rng = np.random.default_rng(77)
    n, tsamp = 30_011, 1e-3
    x = (rng.standard_normal(n) * 3 + np.linspace(0, 50, n)).astype(np.float32)
    got = oracle.deredden(x, tsamp, 1.0, 51)
    rmed = oracle.fast_running_median(x, 1000, 51)
    assert got.dtype == np.float32
    assert np.array_equal(got, (x.astype(np.float64) - rmed).astype(np.float32))
    # a linear trend is removed away from the ends (first/last half block)
    sf = int(1000 / 51)
    mid = slice(sf, n - 2 * sf)
    assert abs(np.mean(got[mid])) < 0.1


def test_normalise_moments(oracle):
    x = (np.random.default_rng(0).standard_normal(100_000) * 5 + 3).astype(np.float32)
    y = oracle.normalise(x)
    assert abs(y.mean()) < 1e-6 and abs(y.std() - 1) < 1e-6


# ---------------------------------------------------------------------------
# Golden fixtures written from the compiled reference (tests/golden/make_golden.py)
# ---------------------------------------------------------------------------

def _golden():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.npz")
    return np.load(path)


def test_oracle_matches_golden_fixtures():
    from oracle import oracle as O
    G = _golden()
    L = O.oracle()
    for i in range(5):
        assert np.array_equal(L.ffa2(G[f"ffa2_{i}_in"]), G[f"ffa2_{i}_out"])
    assert np.array_equal(L.snr2(G["snr2_in"], G["snr2_widths"], 1.3), G["snr2_out"])
    for i in range(4):
        got = L.downsample(G["downsample_in"], float(G[f"downsample_{i}_f"]))
        assert np.array_equal(got, G[f"downsample_{i}_out"])
    for i in range(2):
        got = L.running_median(G["rmed_in"], int(G[f"rmed_{i}_width"]))
        assert np.array_equal(got, G[f"rmed_{i}_out"])
    a = G["pgram_args"]
    P, F, S = L.periodogram(G["pgram_in"], float(G["pgram_tsamp"]), G["pgram_widths"],
                            a[0], a[1], int(a[2]), int(a[3]))
    assert np.array_equal(P, G["pgram_periods"])
    assert np.array_equal(F, G["pgram_foldbins"])
    assert np.array_equal(S, G["pgram_snrs"])


def test_periodogram_f64_recomputation(oracle):
    """The float64 recomputation used for error statements agrees with the
    float32 oracle to float32 accuracy and with the independent float64 torch
    reimplementation of the path (tests/fp64_periodogram.py, all transforms)
    to float64 accuracy, apart from the float32 rounding of the downsampling
    edge weights that the latter keeps (~1e-8 relative)."""
    import torch
    from tests import fp64_periodogram as R
    rng = np.random.default_rng(21)
    N, tsamp, pr = 20_000, 1e-3, (0.3, 0.6, 30, 33)
    x = oracle.normalise(rng.standard_normal(N).astype(np.float32) + 0.5 * np.sin(np.arange(N)))
    w = oracle.generate_width_trials(30)
    S64 = oracle.oracle().periodogram_f64(x, tsamp, w, *pr)
    _, _, S32 = oracle.oracle().periodogram(x, tsamp, w, *pr)
    assert S64.shape == S32.shape and S64.dtype == np.float64
    np.testing.assert_allclose(S32, S64, rtol=1e-5, atol=1e-5)
    rows = R.periodogram_rows(torch.tensor(x, dtype=torch.float64), tsamp, [int(v) for v in w],
                              *pr, keep=lambda t: True, normalise_input=False)
    T = torch.cat([rows[k] for k in sorted(rows)]).numpy()
    np.testing.assert_allclose(S64, T, rtol=1e-6, atol=1e-6)
    # the float32 path's error is of float32 size, the float64 one's far below
    assert np.abs(S32 - S64).max() > 10 * np.abs(S64 - T).max() or np.abs(S32 - S64).max() < 1e-6
