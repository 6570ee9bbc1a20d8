"""FIR whitening front end (rt_fir_forward / rt_fir_grad_taps / rt_adam_update)
against float64 numpy restatements, the adjoint (dot-product) identities of
both gradients, and the whitening optimiser on a small batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np_fir(x, h):
    """y[n] = sum_k h[k] x[n + k - c], zero outside the series (float64)."""
    c = (h.size - 1) // 2
    xp = np.concatenate([np.zeros(c), x.astype(np.float64), np.zeros(c)])
    return np.correlate(xp, h.astype(np.float64), mode="valid")


def _np_grad_taps(x, g, T):
    c = (T - 1) // 2
    xp = np.concatenate([np.zeros(c), x.astype(np.float64), np.zeros(c)])
    return np.array([np.dot(g.astype(np.float64), xp[k:k + x.size]) for k in range(T)])


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


@pytest.mark.parametrize("S,N,T", [(1, 1, 1), (3, 5000, 129), (2, 20000, 31), (2, 777, 1023)])
def test_fir_forward_and_reverse(torch, S, N, T):
    from riptide_amd import _lib
    from riptide_amd.fir import fir_forward
    rng = np.random.default_rng(N + T)
    x = rng.standard_normal((S, N)).astype(np.float32)
    h = rng.standard_normal((S, T)).astype(np.float32) / np.sqrt(T)
    y = _lib.to_numpy(fir_forward(_lib.to_device(x), _lib.to_device(h)))
    yr = _lib.to_numpy(fir_forward(_lib.to_device(x), _lib.to_device(h), reverse=True))
    for s in range(S):
        np.testing.assert_allclose(y[s], _np_fir(x[s], h[s]), rtol=1e-5, atol=2e-5)
        np.testing.assert_allclose(yr[s], _np_fir(x[s], h[s][::-1]), rtol=1e-5, atol=2e-5)


def test_fir_identity_taps(torch):
    from riptide_amd import _lib
    from riptide_amd.fir import fir_forward, identity_taps
    x = np.random.default_rng(0).standard_normal((4, 10000)).astype(np.float32)
    y = fir_forward(_lib.to_device(x), identity_taps(4, 129))
    assert np.array_equal(_lib.to_numpy(y), x)


@pytest.mark.parametrize("S,N,T", [(2, 9000, 129), (1, 4096 * 3 + 5, 9)])
def test_fir_grad_taps(torch, S, N, T):
    from riptide_amd import _lib
    from riptide_amd.fir import fir_grad_taps
    rng = np.random.default_rng(S * N)
    x = rng.standard_normal((S, N)).astype(np.float32)
    g = rng.standard_normal((S, N)).astype(np.float32)
    gt = _lib.to_numpy(fir_grad_taps(_lib.to_device(x), _lib.to_device(g), T))
    for s in range(S):
        want = _np_grad_taps(x[s], g[s], T)
        np.testing.assert_allclose(gt[s], want, rtol=1e-4, atol=1e-3)


def test_fir_adjoint_identities(torch):
    """<g, F(x, h)> = <F^T g, x> = <dF/dh^T g, h>: both hand-written
    adjoints of the (bilinear) filter, through torch.autograd."""
    from riptide_amd import _lib
    from riptide_amd.fir import fir_filter
    rng = np.random.default_rng(7)
    S, N, T = 3, 6000, 65
    x = _lib.to_device(rng.standard_normal((S, N)).astype(np.float32)).requires_grad_(True)
    h = _lib.to_device(rng.standard_normal((S, T)).astype(np.float32)).requires_grad_(True)
    g = _lib.to_device(rng.standard_normal((S, N)).astype(np.float32))
    y = fir_filter(x, h)
    (y * g).sum().backward()
    lhs = float((y.detach().double() * g.double()).sum())
    via_x = float((x.grad.double() * x.detach().double()).sum())
    via_h = float((h.grad.double() * h.detach().double()).sum())
    assert abs(via_x - lhs) < 1e-4 * (abs(lhs) + np.sqrt(S * N))
    assert abs(via_h - lhs) < 1e-4 * (abs(lhs) + np.sqrt(S * N))


def test_adam_update_matches_formula(torch):
    from riptide_amd import _lib
    from riptide_amd.fir import Adam
    rng = np.random.default_rng(3)
    p0 = rng.standard_normal(1000).astype(np.float32)
    p = _lib.to_device(p0)
    opt = Adam(p, lr=1e-2, betas=(0.9, 0.99), eps=1e-8)
    pr, m, v = p0.astype(np.float64), np.zeros(1000), np.zeros(1000)
    for t in range(1, 6):
        g = rng.standard_normal(1000).astype(np.float32)
        opt.step(_lib.to_device(g))
        m = 0.9 * m + 0.1 * g
        v = 0.99 * v + 0.01 * g.astype(np.float64) ** 2
        pr = pr - 1e-2 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.99 ** t)) + 1e-8)
    np.testing.assert_allclose(_lib.to_numpy(p), pr, rtol=1e-5, atol=1e-6)


def test_whitening_gradient_direction_and_steps(torch):
    """The tap gradient of the whitening objective points uphill (central
    difference along it, float32 path), and a few Adam steps raise the
    objective of a red-noise batch with pulses."""
    from riptide_amd import _lib
    from riptide_amd.autograd import periodogram_fn, periodogram_grid
    from riptide_amd.ffautils import generate_width_trials
    from riptide_amd.fir import (fir_filter, identity_taps, near_period_rows,
                                 optimise_whitening, whitening_objective)
    rng = np.random.default_rng(11)
    S, N, tsamp, T = 2, 2 ** 16, 1e-3, 33
    t = np.arange(N) * tsamp
    periods = [0.73, 1.21]
    xs = []
    for s in range(S):
        ph = (t / periods[s]) % 1.0
        pulse = np.exp(-0.5 * ((ph - 0.5) / 0.02) ** 2)
        red = np.cumsum(rng.standard_normal(N)) * 0.05
        xs.append((rng.standard_normal(N) + red + 0.4 * pulse).astype(np.float32))
    X = _lib.to_device(np.stack(xs))
    kw = dict(period_min=0.5, period_max=2.0, bins_min=240, bins_max=260)
    widths = generate_width_trials(240)
    grid, _ = periodogram_grid(N, tsamp, 0.5, 2.0, 240, 260)
    rows = near_period_rows(grid, periods[0])

    def objective(taps):
        y = fir_filter(X[:1], taps)
        return whitening_objective(periodogram_fn(y[0], tsamp, widths, **kw), rows)

    h = identity_taps(1, T).requires_grad_(True)
    objective(h).backward()
    G = h.grad.detach()
    d = G / G.norm()
    eps = 2e-3
    up = float(objective(identity_taps(1, T) + eps * d))
    dn = float(objective(identity_taps(1, T) - eps * d))
    fd = (up - dn) / (2 * eps)
    assert fd > 0 and abs(fd - float(G.norm())) < 0.3 * float(G.norm())

    taps, hist = optimise_whitening(X, tsamp, periods, steps=4, ntaps=T, lr=5e-3, **kw)
    assert all(np.all(np.isfinite(hh)) for hh in hist)
    assert hist[-1].mean() > hist[0].mean()
