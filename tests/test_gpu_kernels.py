"""GPU parity of the kernel-level API against the CPU oracle (same seeded
inputs): ffa2 and the running median bit-exact; S/N and downsampling within
the float32 tolerances stated per test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    from riptide_amd import libcpp
    return libcpp


FFA_SHAPES = [
    (1, 7), (2, 9), (3, 5), (8, 8), (7, 240), (37, 260), (257, 5), (536, 250),
    (150, 260),     # one workgroup, whole transform
    (2684, 260),    # leaf pass + one cone pass
    (1342, 520),    # leaf pass + two cone passes
    (5368, 240),    # deepest tree of the search configurations
    (8195, 5),      # first row count where the float32 row map != integer rounding
    (40, 3000),     # rows too wide to tile: global-memory levels
    (5, 20000),
]


@pytest.mark.parametrize("shape", FFA_SHAPES)
def test_ffa2_bit_exact(lc, oracle, shape):
    rng = np.random.default_rng(hash(shape) % 2**32)
    x = rng.standard_normal(shape).astype(np.float32)
    got = lc.ffa2(x)
    want = oracle.oracle().ffa2(x)
    assert got.dtype == np.float32 and got.shape == x.shape
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


@pytest.mark.parametrize("wg_shape", [0, 1, 2])
@pytest.mark.parametrize("shape", [(150, 260), (2684, 260), (1342, 520), (97, 1000)])
def test_ffa2_workgroup_shapes(lc, oracle, wg_shape, shape):
    """Every workgroup shape (threads, LDS budget, register staging) tiles the
    same transform bit-exactly."""
    from riptide_amd import _lib
    rng = np.random.default_rng(shape[0] + wg_shape)
    x = rng.standard_normal(shape).astype(np.float32)
    _lib.call("rt_set_ffa_shape", wg_shape)
    try:
        got = lc.ffa2(x)
    finally:
        _lib.call("rt_set_ffa_shape", -1)
    assert np.array_equal(got, oracle.oracle().ffa2(x))


# Row lengths of every FFA kernel variant (columns per lane k = ceil(bins/64):
# variants hold k in {2v, 2v+1} up to k = 17, then one variant for wider rows),
# with more rows than one workgroup tiles so that cone passes run too.
VARIANT_BINS = [64, 100, 190, 300, 400, 470, 600, 700, 800, 900, 1024, 1050, 1100, 1500, 2500]


@pytest.mark.parametrize("wg_shape", [0, 1, 2])
@pytest.mark.parametrize("bins", VARIANT_BINS)
def test_ffa2_every_kernel_variant(lc, oracle, wg_shape, bins):
    from riptide_amd import _lib
    rows = max(60, 120000 // bins)
    rng = np.random.default_rng(bins * 3 + wg_shape)
    x = rng.standard_normal((rows, bins)).astype(np.float32)
    _lib.call("rt_set_ffa_shape", wg_shape)
    try:
        got = lc.ffa2(x)
    finally:
        _lib.call("rt_set_ffa_shape", -1)
    assert np.array_equal(got, oracle.oracle().ffa2(x))


def test_ffa2_kat_8x8(lc):
    x = np.zeros((8, 8), np.float32)
    x[:, 7] = 1
    out = lc.ffa2(x)
    assert np.array_equal(out[0], [0, 0, 0, 0, 0, 0, 0, 8])
    assert np.array_equal(out[7], [1] * 8)


@pytest.mark.parametrize("rows,cols", [(1, 64), (50, 260), (300, 240), (7, 520), (3, 1500)])
def test_snr2_parity(lc, oracle, rows, cols):
    rng = np.random.default_rng(rows * 1000 + cols)
    x = (rng.standard_normal((rows, cols)) * 5).astype(np.float32)
    w = oracle.generate_width_trials(min(cols - 1, 480))
    w = w[w < cols]
    got = lc.snr2(x, w, 1.7)
    want = oracle.oracle().snr2(x, w, 1.7)
    # S/N tolerance of the search: 1e-5 relative and 1e-4 absolute
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)
    # the float-float scan reproduces the float64 prefix sums: nearly all bits agree
    assert np.mean(got == want) > 0.99


@pytest.mark.parametrize("f", [1.5, 2.0, 6.510416666666667, 13.777, 600.3, 4999.5])
def test_downsample_parity(lc, oracle, f):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(200003).astype(np.float32)
    got = lc.downsample(x, f)
    want = oracle.oracle().downsample(x, f)
    assert got.shape == want.shape
    # same index bounds and edge weights; only the float32 summation order
    # differs from the vectorised reference loop
    # float32 rounding of ~f sequential adds of O(sqrt(f)) partial sums
    np.testing.assert_allclose(got, want, rtol=0, atol=3e-7 * f + 1e-6)


@pytest.mark.parametrize("width", [1, 3, 5, 7, 11, 25, 37, 101])
def test_running_median_bit_exact(lc, oracle, width):
    rng = np.random.default_rng(width)
    x = rng.standard_normal(5000).astype(np.float32)
    assert np.array_equal(lc.running_median(x, width), oracle.oracle().running_median(x, width))


def test_helpers(lc, oracle):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(37).astype(np.float32)
    y = rng.standard_normal(37).astype(np.float32)
    for shift in (0, 1, 5, 36, 37, 100):
        assert np.array_equal(lc.rollback(x, shift), np.roll(x, -shift))
        assert np.array_equal(lc.fused_rollback_add(x, y, shift), x + np.roll(y, -shift))
    for nsum in (10, 37, 50, 111):
        got = lc.circular_prefix_sum(x, nsum)
        want = oracle.oracle().circular_prefix_sum(x, nsum)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)


def test_benchmark_ffa2(lc):
    t = lc.benchmark_ffa2(64, 256, 3)
    assert 0 < t < 1.0


def test_golden_fixtures(lc):
    """The HIP path against the compiled reference's own outputs
    (tests/golden/reference_outputs.npz, written by tests/golden/make_golden.py)."""
    import os
    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.npz"))
    for i in range(5):
        assert np.array_equal(lc.ffa2(G[f"ffa2_{i}_in"]), G[f"ffa2_{i}_out"])
    np.testing.assert_allclose(lc.snr2(G["snr2_in"], G["snr2_widths"], 1.3), G["snr2_out"],
                               rtol=1e-5, atol=1e-4)
    for i in range(4):
        f = float(G[f"downsample_{i}_f"])
        got = lc.downsample(G["downsample_in"], f)
        np.testing.assert_allclose(got, G[f"downsample_{i}_out"], rtol=0, atol=3e-7 * f + 1e-6)
    for i in range(2):
        got = lc.running_median(G["rmed_in"], int(G[f"rmed_{i}_width"]))
        assert np.array_equal(got, G[f"rmed_{i}_out"])
    a = G["pgram_args"]
    P, F, S = lc.periodogram(G["pgram_in"], float(G["pgram_tsamp"]), G["pgram_widths"],
                             a[0], a[1], int(a[2]), int(a[3]))
    assert np.array_equal(P, G["pgram_periods"])
    assert np.array_equal(F, G["pgram_foldbins"])
    np.testing.assert_allclose(S, G["pgram_snrs"], rtol=1e-5, atol=1e-4)
