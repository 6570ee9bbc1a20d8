"""Batched differentiable periodogram (rt_grad_batch_*, BatchPeriodogram):
normalisation and backward bit-identical to the one-series entry points on
every series, the batched forward equal to the one-series periodogram within
the S/N tolerance, and directional derivatives of the batched whitening
objective at the C5 series length (2^21 samples, 129 taps)."""
import ctypes

import numpy as np
import pytest

from riptide_amd import synthetic

pytestmark = pytest.mark.gpu

S, N, TSAMP = 5, 2 ** 17, 256e-6
RANGE = (0.5, 2.0, 240, 260)


def _batch(seed=3):
    import torch
    from riptide_amd import _lib
    rng = np.random.default_rng(seed)
    xs = [synthetic.series(N, TSAMP, seed=seed + k, red_index=-2.0,
                           pulses=[(float(rng.uniform(0.6, 1.8)), 0.03, 12.0)]) * (1 + k)
          for k in range(S)]
    return torch.stack([_lib.to_device(x) for x in xs])


def test_batch_normalise_matches_one_series():
    import torch
    from riptide_amd import _lib
    from riptide_amd.autograd import BatchPeriodogram
    X = _batch()
    bp = BatchPeriodogram(S, N, TSAMP, None, *RANGE)
    try:
        _, Y, inv_std = bp.forward_nograd(X)
        for k in range(S):
            y = torch.empty_like(X[k])
            stats = np.zeros(4, np.float32)
            _lib.call("rt_normalise_ex", _lib.ptr(X[k]), N, _lib.ptr(y), _lib.host_ptr(stats),
                      _lib.stream())
            assert torch.equal(Y[k], y)
            assert inv_std[k] == np.float32(1.0 / (np.float64(stats[2]) + np.float64(stats[3])))
    finally:
        bp.close()


def _one_series_backward(y, inv_std, gsnr, widths):
    import torch
    from riptide_amd import _lib
    gy = torch.empty_like(y)
    w = np.ascontiguousarray(widths, dtype=np.uint64)
    _lib.call("rt_periodogram_backward", _lib.ptr(y), N, TSAMP, _lib.host_ptr(w), w.size,
              *RANGE, _lib.ptr(gsnr.contiguous()), _lib.ptr(gy), _lib.stream())
    gx = torch.empty_like(y)
    _lib.call("rt_normalise_backward", _lib.ptr(gy), _lib.ptr(y), N, float(inv_std),
              _lib.ptr(gx), _lib.stream())
    return gx


@pytest.mark.parametrize("kind", ["sparse", "dense"])
def test_batch_backward_bit_identical(kind):
    """Sparse: a few rows per width per series (what max/median objectives
    give); dense: every row of a few whole transforms of some series and none
    of others (series with no gradient get zeros)."""
    import torch
    from riptide_amd.autograd import BatchPeriodogram
    X = _batch(seed=11)
    bp = BatchPeriodogram(S, N, TSAMP, None, *RANGE)
    try:
        _, Y, inv_std = bp.forward_nograd(X)
        L, nw = bp.length, bp.widths.size
        rng = np.random.default_rng(5 if kind == "sparse" else 6)
        G = np.zeros((S, L, nw), np.float32)
        for k in range(S):
            if kind == "sparse":
                for w in range(nw):
                    G[k, rng.integers(0, L, 2), w] = rng.standard_normal(2)
            elif k != 2:
                a = int(rng.integers(0, L - 3000))
                G[k, a:a + 3000] = rng.standard_normal((3000, nw))
        Gd = torch.as_tensor(G, device=X.device)
        gx = bp.backward_raw(Y, inv_std, Gd)
        for k in range(S):
            want = _one_series_backward(Y[k].contiguous(), inv_std[k], Gd[k], bp.widths)
            assert torch.equal(gx[k], want), k
        if kind == "dense":
            assert not torch.any(gx[2])
    finally:
        bp.close()


def test_batch_forward_matches_periodogram_fn():
    import torch
    from riptide_amd.autograd import BatchPeriodogram, periodogram_fn
    X = _batch(seed=21)
    bp = BatchPeriodogram(S, N, TSAMP, None, *RANGE)
    try:
        snrs = bp(X)
        torch.cuda.synchronize()
        bp.batch.check()
        for k in range(S):
            want = periodogram_fn(X[k].contiguous(), TSAMP, bp.widths, *RANGE)
            np.testing.assert_allclose(snrs[k].cpu().numpy(), want.cpu().numpy(), rtol=1e-5,
                                       atol=1e-4)
    finally:
        bp.close()


@pytest.mark.timeout(600)
def test_batch_whitening_objective_directional_derivatives():
    """Four series of 2^21 samples at 128 us (red noise of index -2, pulses
    at 0.5-5 s), 129-tap filters, search 0.5-5 s / 240-260 bins: the
    derivative of the summed whitening objective along the tap gradient and
    along random directions (float32 central differences) agrees with the
    batched adjoint's gradient."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.autograd import BatchPeriodogram
    from riptide_amd.fir import (_row_index, batch_whitening_objective, fir_filter,
                                 identity_taps)
    dev = _lib.device()
    nser, n, tsamp, T = 4, 2 ** 21, 128e-6, 129
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    periods = [0.61, 1.37, 2.9, 4.4]
    X = torch.stack([synthetic.series_torch(torch, n, tsamp, g, dev, red_index=-2.0,
                                            pulses=[(p, 0.02, 15.0)]) for p in periods])
    bp = BatchPeriodogram(nser, n, tsamp, None, 0.5, 5.0, 240, 260)
    try:
        grid, _ = bp.grid()
        rows = _row_index(grid, periods)

        def objective(taps):
            return batch_whitening_objective(bp(fir_filter(X, taps)), rows).sum()

        h = identity_taps(nser, T).requires_grad_(True)
        objective(h).backward()
        G = h.grad.detach()
        assert torch.isfinite(G).all() and float(G.norm()) > 0
        rng = np.random.default_rng(8)
        eps = 2e-3
        with torch.no_grad():
            dirs = [G / G.norm()] + [torch.as_tensor(rng.standard_normal((nser, T)),
                                                     dtype=torch.float32, device=dev)
                                     for _ in range(2)]
            for d in dirs:
                d = d / d.norm()
                up = float(objective(identity_taps(nser, T) + eps * d))
                dn = float(objective(identity_taps(nser, T) - eps * d))
                fd = (up - dn) / (2 * eps)
                proj = float((G * d).sum())
                assert abs(fd - proj) < 0.3 * float(G.norm()), (fd, proj)
    finally:
        bp.close()
