"""Peak detection in periodograms (peak_detection.py:38-240 of the reference)
with the S/N array left on the device.

Per width trial the reference (1) cuts the periodogram into m segments of
p = L // m consecutive trials spanning segwidth / T in frequency, takes the
median frequency and the 25/50/75th S/N percentiles of each, (2) fits a
degree-polydeg polynomial in log f to the control points median + nstd * IQR
/ 1.349 (a constant smin below minseg segments), (3) selects the trials whose
S/N exceeds both smin and the polynomial, (4) clusters the selected
frequencies within clrad / T and keeps each cluster's best trial.

Here the order statistics of (1) and the selection of (3) run on the GPU
(rt_segment_select: exact radix select; rt_peak_candidates: float32
threshold with an error bound, the trials inside the bound decided on the
host in float64 exactly as the reference evaluates it).  The percentile
interpolation, the polynomial fit and the clustering of the few selected
trials run on the host with numpy's own formulas, so the peaks and the
polynomial coefficients are the reference's.
"""
import typing
from math import ceil

import numpy as np

from . import _lib
from .clustering import cluster1d

_QUANTILES = (0.25, 0.5, 0.75)


class Peak(typing.NamedTuple):
    """Parameters of a peak found in a Periodogram (peak_detection.py:14-37)."""

    period: float
    freq: float
    width: int
    ducy: float
    iw: int
    ip: int
    snr: float
    dm: float

    def summary_dict(self):
        attrs = ("period", "freq", "dm", "width", "ducy", "snr")
        return {a: getattr(self, a) for a in attrs}


def _torch():
    import torch
    return torch


def percentile_ranks(p):
    """(lower, upper, gamma) of numpy's linear percentile interpolation of the
    25th, 50th and 75th percentiles of p values (virtual index (p-1) * q)."""
    out = []
    for q in _QUANTILES:
        vi = (p - 1) * q
        lo = int(np.floor(vi))
        if vi >= p - 1:
            lo = hi = p - 1
        else:
            hi = lo + 1
        out.append((lo, hi, vi - lo))
    return out


def lerp(a, b, gamma):
    """numpy's linear interpolation between order statistics (float64):
    a + (b - a) * g below g = 0.5, b - (b - a) * (1 - g) from it."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = b - a
    return np.where(gamma >= 0.5, b - d * (1.0 - gamma), a + d * gamma)


def _smin_lo(smin):
    """Largest float32 <= smin: for a float32 s, s > smin <=> s > smin_lo."""
    v = np.float32(smin)
    if float(v) > smin:
        v = np.nextafter(v, np.float32(-np.inf))
    return float(v)


def _horner(coeffs, x):
    """poly1d evaluation as numpy does it (y = y * x + c, float64)."""
    y = 0.0
    for c in coeffs:
        y = y * x + float(c)
    return y


class PeakFinder:
    """find_peaks for every periodogram of one plan (period grid): the
    segment layout, log-frequencies and segment median frequencies are
    computed once."""

    def __init__(self, periods, foldbins, widths, tobs, segwidth=5.0):
        self.periods = np.asarray(periods, np.float64)
        self.freqs = 1.0 / self.periods
        self.foldbins = np.asarray(foldbins)
        self.widths = np.asarray(widths)
        self.tobs = float(tobs)
        self.segwidth = float(segwidth)
        f = self.freqs
        w = self.segwidth / self.tobs
        self.m = ceil(abs(f[-1] - f[0]) / w)
        self.p = len(f) // self.m
        n = self.m * self.p
        self.fc = np.median(f[:n].reshape(self.m, self.p), axis=1)
        self.logf = np.log(f)
        self._dev = None

    def _device_arrays(self):
        if self._dev is None:
            torch = _torch()
            dev = _lib.device()
            self._dev = torch.tensor(self.logf.astype(np.float32), device=dev)
        return self._dev

    def segment_stats(self, snrs):
        """(fc, smed, sstd) per width, as segment_stats (peak_detection.py:
        40-89): arrays of shape (m, nw)."""
        torch = _torch()
        nw = snrs.shape[1]
        pr = percentile_ranks(self.p)
        ranks = []
        for lo, hi, _ in pr:
            ranks += [lo, hi]
        ranks = np.asarray(ranks, np.int32)
        out = torch.empty((self.m, nw, len(ranks)), dtype=torch.float32, device=snrs.device)
        _lib.call("rt_segment_select", _lib.ptr(snrs), nw, self.m, self.p, _lib.host_ptr(ranks),
                  len(ranks), _lib.ptr(out), _lib.stream())
        v = out.cpu().numpy().astype(np.float64)
        s25, smed, s75 = (lerp(v[:, :, 2 * k], v[:, :, 2 * k + 1], pr[k][2]) for k in range(3))
        sstd = (s75 - s25) / 1.349
        return self.fc, smed, sstd

    def find(self, snrs, dm=None, smin=6.0, nstd=6.0, minseg=10, polydeg=2, clrad=0.1):
        """(peaks, polycos) of one periodogram whose S/N (num_periods x
        num_widths float32) is a CUDA tensor; as find_peaks
        (peak_detection.py:164-240)."""
        torch = _torch()
        snrs = snrs.contiguous()
        L, nw = snrs.shape
        if L != len(self.freqs):
            raise ValueError("S/N rows do not match the period grid")
        fc, smed, sstd = self.segment_stats(snrs)
        polycos = {}
        coefs = []
        for iw in range(nw):
            sc = smed[:, iw] + nstd * sstd[:, iw]
            if len(fc) >= minseg:
                polyco = np.poly1d(np.polyfit(np.log(fc), sc, polydeg)).coefficients
            else:  # constant threshold below minseg segments, as the reference returns it
                polyco = [smin]
            polycos[iw] = polyco
            coefs.append(np.asarray(polyco, np.float64))
        ncoef = max(len(c) for c in coefs)
        # thresholds of degree > 8 are beyond the device evaluation: the
        # device pass then applies smin only (a threshold far below every
        # S/N) and the host decides every selected value in float64
        host_threshold = ncoef > MAX_DEVICE_COEFS
        if host_threshold:
            ncoef = 1
            C = np.full((nw, 1), -3.0e38, np.float32)
        else:
            C = np.zeros((nw, ncoef), np.float32)
            for iw, c in enumerate(coefs):
                C[iw, ncoef - len(c):] = c  # leading zeros keep the degree
        dev = snrs.device
        logf32 = self._device_arrays()
        d_coef = torch.tensor(C, device=dev)
        counts = torch.zeros(2 * nw, dtype=torch.int32, device=dev)
        cap = L
        sel = torch.empty((nw, cap), dtype=torch.int32, device=dev)
        unc = torch.empty((nw, cap), dtype=torch.int32, device=dev)
        _lib.call("rt_peak_candidates", _lib.ptr(snrs), L, nw, _lib.ptr(logf32), _lib.ptr(d_coef),
                  ncoef, _smin_lo(smin), _lib.ptr(counts), _lib.ptr(sel), _lib.ptr(unc), cap,
                  _lib.stream())
        cnt = counts.cpu().numpy().reshape(nw, 2)
        f = self.freqs
        peaks = []
        for iw, width in enumerate(self.widths):
            ns, nu = int(cnt[iw, 0]), int(cnt[iw, 1])
            sure = sel[iw, :ns].cpu().numpy().astype(np.int64)
            maybe = unc[iw, :nu].cpu().numpy().astype(np.int64)
            if host_threshold:
                sure, maybe = sure[:0], np.concatenate([sure, maybe])
                nu = maybe.size
            if nu:
                s_maybe = snrs[torch.as_tensor(maybe, device=dev), iw].cpu().numpy().astype(np.float64)
                thr = np.array([_horner(polycos[iw], self.logf[i]) for i in maybe])
                maybe = maybe[(s_maybe > thr) & (s_maybe > smin)]
            indices = np.sort(np.concatenate([sure, maybe]))
            if not len(indices):
                continue
            s_sel = snrs[torch.as_tensor(indices, device=dev), iw].cpu().numpy().astype(np.float64)
            for cl in cluster1d(f[indices], clrad / self.tobs):
                ipeak = int(indices[cl][s_sel[cl].argmax()])
                snr = float(s_sel[cl].max())
                peaks.append(Peak(freq=float(f[ipeak]), period=float(1.0 / f[ipeak]),
                                  width=int(width),
                                  ducy=float(float(width) / self.foldbins[ipeak]),
                                  iw=int(iw), ip=ipeak, snr=snr, dm=dm))
        peaks = sorted(peaks, key=lambda pk: pk.snr, reverse=True)
        return peaks, polycos


# Coefficients of the largest threshold polynomial the device pass evaluates
# (degree 8, rt_peak_candidates); larger degrees are decided on the host.
MAX_DEVICE_COEFS = 9


def find_peaks(pgram, smin=6.0, segwidth=5.0, nstd=6.0, minseg=10, polydeg=2, clrad=0.1):
    """Peaks of a Periodogram (peak_detection.py:164-240); pgram.snrs may be
    a numpy array (copied to the device) or a CUDA tensor."""
    finder = PeakFinder(pgram.periods, pgram.foldbins, pgram.widths, pgram.tobs,
                        segwidth=segwidth)
    snrs = getattr(pgram, "snrs_device", None)
    if snrs is None:
        snrs = pgram.snrs
    if isinstance(snrs, np.ndarray):
        snrs = _lib.to_device(np.ascontiguousarray(snrs, np.float32))
    return finder.find(snrs, dm=pgram.metadata["dm"], smin=smin, nstd=nstd, minseg=minseg,
                       polydeg=polydeg, clrad=clrad)
