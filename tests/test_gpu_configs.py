"""Every search configuration run at its configured size on the GPU.

C2  plain ffa_search of one 2^23-sample series at 64 us (white + red noise of
    PSD index -1.5 equal to the white level at 0.1 Hz, pulse trains at
    0.2371 s / 3 % / S/N 12 and 4.9133 s / 1 % / S/N 25), P 0.1-10 s, bins
    480-520: periods bit-exact, S/N within 1e-5 rel / 1e-4 abs of the oracle,
    both pulses found.
C3  a slice of the heterogeneous batch: series whose lengths come from the
    configuration's log-uniform 2^21-2^23 draw (multiples of 1024), both bins
    classes, P 0.1-10 s, through BatchSearch eagerly and replayed from its
    hipGraph, every series against the oracle.
C4  gradient of the windowed-max loss at 2^20 samples against the float64
    reimplementation and 64 float64 central differences.
C5  FIR whitening at 2^21 samples with 129 taps: directional derivatives of
    the whitening objective along the tap gradient, the adjoint identities of
    the filter on the full 128 x 2^21 batch, and the tap gradient against
    float64 numpy on two of its series.
"""
import numpy as np
import pytest

from riptide_amd import synthetic

pytestmark = pytest.mark.gpu

SNR_RTOL, SNR_ATOL = 1e-5, 1e-4


@pytest.mark.timeout(600)
def test_c2_full_search(oracle):
    import riptide_amd as rt
    tsamp = 64e-6
    pulses = [(0.2371, 0.03, 12.0), (4.9133, 0.01, 25.0)]
    x = synthetic.series(2 ** 23, tsamp, seed=2, red_index=-1.5, pulses=pulses)
    kw = dict(period_min=0.1, period_max=10.0, bins_min=480, bins_max=520, rmed_width=4.0)
    ts = rt.TimeSeries(x, tsamp)
    tsdr, pg = rt.ffa_search(ts, fpmin=8, ducy_max=0.2, wtsp=1.5, **kw)
    xo, wo, Po, Fo, So = oracle.ffa_search(x, tsamp, **kw)
    assert pg.snrs.shape == (2_652_094, 12) == So.shape
    assert np.array_equal(pg.widths, wo)
    assert np.array_equal(pg.periods, Po) and np.array_equal(pg.foldbins, Fo)
    np.testing.assert_allclose(tsdr.data, xo, rtol=0, atol=8e-6)
    np.testing.assert_allclose(pg.snrs, So, rtol=SNR_RTOL, atol=SNR_ATOL)
    for period, _, snr in pulses:
        sel = np.abs(pg.periods / period - 1) < 2e-3
        assert pg.snrs[sel].max() > 0.8 * snr
    # the brighter train is the strongest peak of the whole periodogram
    best = pg.snrs.max(axis=1).argmax()
    assert abs(pg.periods[best] / 4.9133 - 1) < 2e-3


@pytest.mark.timeout(900)
def test_c3_slice_batch(oracle):
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch
    tsamp, pmin, pmax = 64e-6, 0.1, 10.0
    rs = np.random.RandomState(303)
    # the configuration's length draw; the slice keeps six series across the range
    lens = (np.exp(rs.uniform(np.log(2 ** 21), np.log(2 ** 23), size=64)).astype(np.int64)
            // 1024 * 1024)
    order = np.argsort(lens)
    pick = order[[0, 12, 25, 38, 51, 63]]
    sizes = [int(lens[i]) for i in pick]
    bmin = np.array([240, 480, 240, 480, 480, 240])
    bmax = np.where(bmin == 240, 260, 520)
    data = []
    for k, n in enumerate(sizes):
        pulses = []
        if k % 2 == 0:
            pulses = [(float(np.exp(rs.uniform(np.log(0.2), np.log(8.0)))),
                       rs.uniform(0.01, 0.05), rs.uniform(8, 30))]
        data.append(synthetic.series(n, tsamp, seed=700 + k, red_index=-1.5, pulses=pulses))
    b = BatchSearch(sizes, tsamp, bmin, bmax, pmin, pmax, rmed_width=4.0, group_size=4)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    want = []
    for i, x in enumerate(data):
        _, w, P, F, S = oracle.ffa_search(x, tsamp, period_min=pmin, period_max=pmax,
                                          bins_min=int(bmin[i]), bins_max=int(bmax[i]),
                                          rmed_width=4.0)
        Pb, Fb = b.grid(i)
        assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
        assert np.array_equal(b.widths[i], w)
        want.append(S)
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        torch.cuda.synchronize()
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i in range(len(data)):
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), want[i], rtol=SNR_RTOL,
                                       atol=SNR_ATOL)
    b.close()


@pytest.mark.timeout(600)
def test_c4_gradient_full_size():
    import torch
    from riptide_amd import _lib
    from riptide_amd.autograd import periodogram_fn, periodogram_grid
    from tests import fp64_periodogram as R
    from tests.test_gpu_autograd import _c4_like, _fp64_loss
    n, tsamp, p0, frac = 2 ** 20, 256e-6, 1.3711, 0.005
    pr = (0.5, 2.0, 240, 260)
    x = _c4_like(n, tsamp, p0, seed=4)
    w = R.generate_width_trials(240)
    xg = torch.tensor(x, device=_lib.device(), requires_grad=True)
    snrs = periodogram_fn(xg, tsamp, w, *pr)
    P, _ = periodogram_grid(n, tsamp, *pr)
    mask = torch.as_tensor(np.abs(P / p0 - 1) <= frac, device=xg.device)
    loss = snrs[mask].max(dim=0).values.sum()
    loss.backward()
    g_gpu = xg.grad.double().cpu().numpy()
    x64 = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    L64 = _fp64_loss(x64, tsamp, w, pr, p0, frac)
    L64.backward()
    g_ref = x64.grad.numpy()
    loss_gpu, loss_ref = float(loss.detach()), float(L64.detach())
    assert abs(loss_gpu - loss_ref) < 1e-4 * abs(loss_ref)
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-4, rel
    rng = np.random.default_rng(64)
    h = 1e-3
    errs = []
    with torch.no_grad():
        xd = torch.tensor(x, dtype=torch.float64)
        for _ in range(64):
            v = rng.standard_normal(n)
            v /= np.linalg.norm(v)
            vt = torch.tensor(v)
            fd = (float(_fp64_loss(xd + h * vt, tsamp, w, pr, p0, frac))
                  - float(_fp64_loss(xd - h * vt, tsamp, w, pr, p0, frac))) / (2 * h)
            errs.append(abs(fd - float(g_gpu @ v)) / max(abs(fd), 1e-3))
    errs = np.asarray(errs)
    assert np.median(errs) < 1e-3 and np.mean(errs < 1e-2) >= 0.9, errs


def _c5_batch(torch, dev, nseries, nsamp=2 ** 21, tsamp=128e-6, seed=55):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rs = np.random.RandomState(seed)
    periods = np.exp(rs.uniform(np.log(0.5), np.log(5.0), size=nseries))
    nfft = 1 << int(np.ceil(np.log2(nsamp)))
    X = torch.empty((nseries, nsamp), device=dev)
    for s in range(nseries):
        X[s] = synthetic.series_torch(torch, nsamp, tsamp, g, dev, red_index=-2.0,
                                      pulses=[(float(periods[s]), 0.02, 15.0)], nfft=nfft)
    return X, periods


@pytest.mark.timeout(600)
def test_c5_tap_gradient_full_length():
    """Whitening objective of 2^21-sample series through a 129-tap filter:
    the directional derivative along the tap gradient (float32 central
    differences) agrees with the gradient's norm, and along random
    directions with the gradient's projection."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.autograd import periodogram_fn, periodogram_grid
    from riptide_amd.ffautils import generate_width_trials
    from riptide_amd.fir import fir_filter, identity_taps, near_period_rows, whitening_objective
    dev = _lib.device()
    nsamp, tsamp, T = 2 ** 21, 128e-6, 129
    X, periods = _c5_batch(torch, dev, 3)
    kw = dict(period_min=0.5, period_max=5.0, bins_min=240, bins_max=260)
    widths = generate_width_trials(240)
    grid, _ = periodogram_grid(nsamp, tsamp, 0.5, 5.0, 240, 260)
    rng = np.random.default_rng(5)
    for s in range(X.shape[0]):
        rows = near_period_rows(grid, periods[s])

        def objective(taps):
            y = fir_filter(X[s:s + 1], taps)
            return whitening_objective(periodogram_fn(y[0], tsamp, widths, **kw), rows)

        h = identity_taps(1, T).requires_grad_(True)
        objective(h).backward()
        G = h.grad.detach()
        assert torch.isfinite(G).all() and float(G.norm()) > 0
        eps = 2e-3
        with torch.no_grad():
            for d in [G / G.norm()] + [torch.as_tensor(rng.standard_normal((1, T)),
                                                       dtype=torch.float32, device=dev)
                                       for _ in range(2)]:
                d = d / d.norm()
                up = float(objective(identity_taps(1, T) + eps * d))
                dn = float(objective(identity_taps(1, T) - eps * d))
                fd = (up - dn) / (2 * eps)
                proj = float((G * d).sum())
                assert abs(fd - proj) < 0.3 * float(G.norm()), (s, fd, proj)


@pytest.mark.timeout(600)
def test_c5_fir_batch_properties():
    """The full C5 filter batch (128 x 2^21, 129 taps): identity taps return
    the input exactly; <g, F(x, h)> = <F^T g, x> = <dF/dh^T g, h>; the tap
    gradient of two series matches float64 numpy."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.fir import fir_filter, fir_forward, fir_grad_taps, identity_taps
    dev = _lib.device()
    S, N, T = 128, 2 ** 21, 129
    X, _ = _c5_batch(torch, dev, S)
    assert torch.equal(fir_forward(X, identity_taps(S, T)), X)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    H = torch.randn((S, T), generator=g, device=dev) / np.sqrt(T)
    Gy = torch.randn((S, N), generator=g, device=dev)
    x = X.clone().requires_grad_(True)
    h = H.clone().requires_grad_(True)
    y = fir_filter(x, h)
    (y * Gy).sum().backward()
    lhs = float((y.detach().double() * Gy.double()).sum())
    via_x = float((x.grad.double() * X.double()).sum())
    via_h = float((h.grad.double() * H.double()).sum())
    scale = abs(lhs) + np.sqrt(S * N) * float(X.abs().max())
    assert abs(via_x - lhs) < 1e-5 * scale and abs(via_h - lhs) < 1e-5 * scale
    gt = fir_grad_taps(X, Gy, T)
    c = (T - 1) // 2
    for s in (0, S - 1):
        xs = X[s].double().cpu().numpy()
        gs = Gy[s].double().cpu().numpy()
        xp = np.concatenate([np.zeros(c), xs, np.zeros(c)])
        want = np.array([np.dot(gs, xp[k:k + N]) for k in range(T)])
        got = gt[s].double().cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())
