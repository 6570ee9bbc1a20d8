"""rt_periodogram_noalloc (every intermediate in a static device array, no
device allocation) against rt_periodogram on the same inputs: identical
periods, fold bins and S/N, on calls in a row of decreasing size (the array
keeps the previous call's contents), with no drop in free device memory, and
a clear error for a search whose intermediates exceed the array."""
import numpy as np
import pytest

from riptide_amd import synthetic

pytestmark = pytest.mark.gpu

CASES = [
    # N, tsamp, pmin, pmax, bmin, bmax
    (2 ** 23, 64e-6, 0.1, 10.0, 480, 520),   # intermediates in several chunks
    (2 ** 20, 256e-6, 0.5, 2.0, 240, 260),   # smaller than the call before
    (2 ** 18, 128e-6, 0.2, 2.0, 480, 520),
]


def _call(name, d_x, case, w, outs):
    from riptide_amd import _lib
    N, tsamp, pmin, pmax, bmin, bmax = case
    P, F, S = outs
    _lib.call(name, _lib.ptr(d_x), N, tsamp, _lib.host_ptr(w), w.size, pmin, pmax, bmin, bmax,
              _lib.ptr(P), _lib.ptr(F), _lib.ptr(S), _lib.stream())


def _outputs(case, nw):
    from riptide_amd import _lib, libcpp
    L = libcpp.periodogram_length(*case)
    return (_lib.empty((L,), "float64"), _lib.empty((L,), "uint32"), _lib.empty((L, nw)))


def test_noalloc_matches_periodogram():
    import torch
    from riptide_amd import _lib
    from riptide_amd.ffautils import generate_width_trials
    dev = _lib.device()
    for k, case in enumerate(CASES):
        N, tsamp, pmin, pmax, bmin, bmax = case
        print(f"noalloc case {k}: N={N} bins {bmin}-{bmax} P {pmin}-{pmax}", flush=True)
        x = synthetic.series(N, tsamp, seed=40 + k, red_index=None,
                             pulses=[(0.37 * (pmin + pmax), 0.03, 15.0)])
        x = (x - x.mean()) / x.std()
        d_x = _lib.to_device(x)
        w = np.ascontiguousarray(generate_width_trials(bmin), dtype=np.uint64)
        want = _outputs(case, w.size)
        _call("rt_periodogram", d_x, case, w, want)
        got = _outputs(case, w.size)
        for t in got:
            t.fill_(7)  # stale values must be overwritten
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info(dev)
        _call("rt_periodogram_noalloc", d_x, case, w, got)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info(dev)
        if k > 0:  # the first call may load the library's device code
            assert free0 - free1 < (4 << 20), (free0, free1)
        for u, v in zip(got, want):
            assert torch.equal(u, v)


def test_noalloc_rejects_oversized_search():
    """2^26 samples at 64 us, P 0.1-10 s, bins 480-520: its packed
    downsampled series alone exceed 256 MiB.  The arrays are real, sized for
    the search, although the call returns before any device work."""
    from riptide_amd import _lib
    from riptide_amd.ffautils import generate_width_trials
    case = (2 ** 26, 64e-6, 0.1, 10.0, 480, 520)
    w = np.ascontiguousarray(generate_width_trials(480), dtype=np.uint64)
    d_x = _lib.empty((case[0],))
    outs = _outputs(case, w.size)
    with pytest.raises(ValueError, match="MiB of scratch, more than the 256 MiB"):
        _call("rt_periodogram_noalloc", d_x, case, w, outs)
