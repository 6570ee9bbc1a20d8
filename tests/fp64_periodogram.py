"""Float64 reimplementation of the periodogram path in torch (CPU), used as
the gradient oracle of the adjoint kernels: normalise -> downsample -> fold ->
FFA -> boxcar S/N, each step differentiable by torch autograd, restricted to
chosen (step, bins) transforms.  Test infrastructure only.

Follows the reference's definitions: time_series.py:67-93 (normalise),
downsample.hpp:44-82 (index bounds and edge weights, computed in float64 as
the reference does), transforms.hpp:13-50 (merge tree and float32 row map),
snr.hpp:37-55 (boxcar S/N), periodogram.hpp:117-201 (plan).
"""
import math

import numpy as np
import torch


def generate_width_trials(nbins, ducy_max=0.20, wtsp=1.5):
    out, w = [], 1
    wmax = int(max(1, ducy_max * nbins))
    while w <= wmax:
        out.append(w)
        w = int(max(w + 1, wtsp * w))
    return out


def plan(size, tsamp, pmin, pmax, bmin, bmax):
    """List of transforms: dict(step f, tau, n, bins, rows, rows_eval, stdnoise, out_row)."""
    ds_ini = pmin / (tsamp * bmin)
    ds_geo = (bmax + 1.0) / bmin
    nds = int(math.ceil(math.log(pmax / pmin) / math.log(ds_geo)))
    trs, row = [], 0
    for ids in range(nds):
        f = ds_ini * ds_geo ** ids
        tau = f * tsamp
        pms = pmax / tau
        n = int(math.floor(size / f))
        k = math.floor(f)
        r = f - k
        x = n * r
        var = f - 1.0 / 3.0 if x > 1 else (k - 1) ** 2 + 2.0 / 3.0 * x ** 2 - x + 1.0
        for bins in range(bmin, min(bmax, n, int(pms)) + 1):
            rows = n // bins
            pc = min(pms, bins + 1.0)
            ce = int(math.ceil(bins * (rows - 1.0) * (1.0 - bins / pc)))
            ev = min(rows, ce)
            trs.append(dict(f=f, tau=tau, n=n, bins=bins, rows=rows, rows_eval=ev,
                            stdnoise=math.sqrt(rows * var), out_row=row))
            row += ev
    return trs


def normalise(x):
    m = x.mean()
    v = ((x - m) ** 2).mean()
    return (x - m) / torch.sqrt(v)


def downsample(x, f):
    N = x.numel()
    n = int(math.floor(N / f))
    k = np.arange(n, dtype=np.float64)
    start = k * f
    end = start + f
    imin = np.floor(start).astype(np.int64)
    imax = np.minimum(np.floor(end), N - 1.0).astype(np.int64)
    wmin = torch.tensor(np.float32((imin + 1) - start).astype(np.float64))
    wmax = torch.tensor(np.float32(end - imax).astype(np.float64))
    C = torch.cat([torch.zeros(1, dtype=x.dtype), torch.cumsum(x, 0)])
    imin_t = torch.as_tensor(imin)
    imax_t = torch.as_tensor(imax)
    interior = torch.where(imax_t > imin_t, C[imax_t] - C[imin_t + 1], torch.zeros_like(wmin))
    return wmin * x[imin_t] + interior + wmax * x[imax_t]


def _row_map(m_child, m, s):
    k = np.float32((np.float32(m_child) - np.float32(1.0)) / (np.float32(m) - np.float32(1.0)))
    return np.float32(np.float64(k) * s + 0.5).astype(np.int64)


def ffa(X):
    m, p = X.shape
    if m == 1:
        return X
    mh = m >> 1
    H = ffa(X[:mh])
    T = ffa(X[mh:])
    s = np.arange(m)
    h = _row_map(mh, m, s)
    t = _row_map(m - mh, m, s)
    cols = (np.arange(p)[None, :] + (s - t)[:, None]) % p
    return H[torch.as_tensor(h)] + torch.gather(T[torch.as_tensor(t)], 1, torch.as_tensor(cols))


def snr(X, widths, stdnoise):
    rows, n = X.shape
    ext = torch.cat([X, X], dim=1)
    C = torch.cat([torch.zeros(rows, 1, dtype=X.dtype), torch.cumsum(ext, 1)], dim=1)
    total = C[:, n]
    out = []
    for w in widths:
        h = math.sqrt((n - w) / (n * w))
        b = w / (n - w) * h
        sums = C[:, w:w + n] - C[:, :n]
        dmax = sums.max(dim=1).values
        out.append(((h + b) * dmax - b * total) / stdnoise)
    return torch.stack(out, dim=1)


def periodogram_rows(x, tsamp, widths, pmin, pmax, bmin, bmax, keep, normalise_input=True):
    """S/N rows of the transforms for which keep(transform) is true, as a dict
    out_row -> tensor (rows_eval, nw); x is a float64 tensor."""
    y = normalise(x) if normalise_input else x
    cache = {}
    out = {}
    for t in plan(x.numel(), tsamp, pmin, pmax, bmin, bmax):
        if t["rows_eval"] <= 0 or not keep(t):
            continue
        if t["f"] not in cache:
            cache[t["f"]] = y if t["f"] == 1 else downsample(y, t["f"])
        d = cache[t["f"]]
        X = d[: t["rows"] * t["bins"]].reshape(t["rows"], t["bins"])
        F = ffa(X)[: t["rows_eval"]]
        out[t["out_row"]] = snr(F, widths, t["stdnoise"])
    return out
