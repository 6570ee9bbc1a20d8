"""Hand-written adjoints against a float64 torch reimplementation of the path
(tests/fp64_periodogram.py, autograd) and float64 central finite differences."""
import numpy as np
import pytest
import torch

from tests import fp64_periodogram as R

pytestmark = pytest.mark.gpu


def _dev():
    from riptide_amd import _lib
    return _lib.device()


def test_snr2_backward_matches_fp64():
    from riptide_amd import _lib
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((40, 260)) * 3).astype(np.float32)
    w = np.array(R.generate_width_trials(240), dtype=np.uint64)
    g = rng.standard_normal((40, w.size)).astype(np.float32)
    dx, dg = _lib.to_device(x), _lib.to_device(g)
    out = torch.empty_like(dx)
    _lib.call("rt_snr2_backward", _lib.ptr(dx), 40, 260, _lib.host_ptr(w), w.size, 1.7,
              _lib.ptr(dg), _lib.ptr(out), _lib.stream())
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    (R.snr(xt, [int(v) for v in w], 1.7) * torch.tensor(g, dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(out.cpu().numpy(), xt.grad.numpy(), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("m,p", [(2, 7), (5, 9), (37, 260), (300, 240), (1342, 50)])
def test_ffa2_backward_is_transpose(m, p):
    from riptide_amd import _lib
    rng = np.random.default_rng(m + p)
    x = rng.standard_normal((m, p)).astype(np.float32)
    y = rng.standard_normal((m, p)).astype(np.float32)
    dx, dy = _lib.to_device(x), _lib.to_device(y)
    ax = torch.empty_like(dx)
    aty = torch.empty_like(dy)
    _lib.call("rt_ffa2", _lib.ptr(dx), m, p, _lib.ptr(ax), _lib.stream())
    _lib.call("rt_ffa2_backward", _lib.ptr(dy), m, p, _lib.ptr(aty), _lib.stream())
    # dot-product test <A x, y> == <x, A^T y>
    lhs = float((ax.double() * dy.double()).sum())
    rhs = float((dx.double() * aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + np.sqrt(m * p))
    # and against float64 autograd of the reimplemented FFA
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    (R.ffa(xt) * torch.tensor(y, dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(aty.cpu().numpy(), xt.grad.numpy(), rtol=1e-5, atol=1e-4)


def _c4_like(n, tsamp, period, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) * tsamp
    ph = (t / period) % 1.0
    sig = np.exp(-0.5 * ((ph - 0.5) / (0.02 / 2.355)) ** 2)
    sig *= 20.0 / np.sqrt((sig ** 2).sum())
    return (rng.standard_normal(n) + sig).astype(np.float32)


def _window_keep(p0, frac):
    def keep(t):
        lo = t["tau"] * t["bins"]
        s_hi = t["rows_eval"] - 1
        hi = t["tau"] * t["bins"] ** 2 / (t["bins"] - s_hi / max(t["rows"] - 1.0, 1.0))
        return hi >= p0 * (1 - frac) and lo <= p0 * (1 + frac)
    return keep


def _fp64_loss(x64, tsamp, w, pr, p0, frac):
    rows = R.periodogram_rows(x64, tsamp, w, *pr, keep=_window_keep(p0, frac))
    sel = []
    for t in R.plan(x64.numel(), tsamp, *pr):
        if t["out_row"] not in rows:
            continue
        s = np.arange(t["rows_eval"])
        per = t["tau"] * t["bins"] ** 2 / (t["bins"] - s / (t["rows"] - 1.0))
        m = np.abs(per / p0 - 1) <= frac
        if m.any():
            sel.append(rows[t["out_row"]][torch.as_tensor(m)])
    return torch.cat(sel).max(dim=0).values.sum()


def test_periodogram_gradient_fp64_and_fd():
    """Config-4 style loss (sum over widths of the max S/N within +-0.5% of
    the injected period) on a 2^18-sample series: the GPU gradient matches the
    float64 autograd gradient, and 64 seeded directional derivatives match
    float64 central differences (h = 1e-3)."""
    from riptide_amd.autograd import periodogram_fn, periodogram_grid
    n, tsamp, p0, frac = 2 ** 18, 256e-6, 1.3711, 0.005
    pr = (0.5, 2.0, 240, 260)
    x = _c4_like(n, tsamp, p0, seed=4)
    w = R.generate_width_trials(240)

    xg = torch.tensor(x, device=_dev(), requires_grad=True)
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
    # the loss is piecewise smooth; allow the odd direction whose +-h step
    # crosses an argmax switch
    assert np.median(errs) < 1e-3 and np.mean(errs < 1e-2) >= 0.9, errs
