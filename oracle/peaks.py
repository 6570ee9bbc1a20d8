"""Oracle of the peak detection (test infrastructure, CPU only): a numpy
restatement of peak_detection.py:40-240 and clustering.py:4-48 of the
reference, with the same numpy calls (np.percentile, np.median, np.polyfit,
np.poly1d, argsort), used to check the device path.  Only tests import it."""
from math import ceil

import numpy as np


def cluster1d(x, r):
    """clustering.py:4-48 (unsorted input)."""
    if not len(x):
        return []
    order = x.argsort()
    cuts = np.where(abs(np.diff(x[order])) > r)[0]
    if not len(cuts):
        return [order]
    edges = np.concatenate(([0], cuts + 1, [len(x)]))
    return [order[a:b] for a, b in zip(edges[:-1], edges[1:])]


def segment_stats(f, s, T, segwidth=5.0):
    """peak_detection.py:40-89."""
    m = ceil(abs(f[-1] - f[0]) / (segwidth / T))
    p = len(f) // m
    n = m * p
    fc = np.median(f[:n].reshape(m, p), axis=1)
    s25, smed, s75 = np.percentile(s[:n].reshape(m, p), (25, 50, 75), axis=-1)
    return fc, smed, (s75 - s25) / 1.349


def find_peaks_single(f, s, T, smin, segwidth, nstd, minseg, polydeg, clrad):
    """peak_detection.py:117-161: (peak indices, polynomial coefficients)."""
    fc, smed, sstd = segment_stats(f, s, T, segwidth)
    sc = smed + nstd * sstd
    if len(fc) >= minseg:
        poly = np.poly1d(np.polyfit(np.log(fc), sc, polydeg))
        polyco = poly.coefficients
    else:
        polyco = [smin]
        poly = np.poly1d(polyco)
    mask = (s > poly(np.log(f))) & (s > smin)
    idx = np.where(mask)[0]
    out = []
    for cl in cluster1d(f[idx], clrad / T):
        ix = idx[cl]
        out.append(ix[s[ix].argmax()])
    return out, polyco


def find_peaks(periods, foldbins, widths, snrs, tobs, dm=None, smin=6.0, segwidth=5.0,
               nstd=6.0, minseg=10, polydeg=2, clrad=0.1):
    """peak_detection.py:164-240 on host arrays: (list of peak tuples
    (period, freq, width, ducy, iw, ip, snr, dm), polycos)."""
    f = 1.0 / np.asarray(periods, np.float64)
    peaks, polycos = [], {}
    for iw, width in enumerate(widths):
        s = snrs[:, iw].astype(float)
        ips, polyco = find_peaks_single(f, s, tobs, smin, segwidth, nstd, minseg, polydeg, clrad)
        for ip in ips:
            peaks.append((float(1.0 / f[ip]), float(f[ip]), int(width),
                          float(float(width) / foldbins[ip]), int(iw), int(ip), float(s[ip]), dm))
        polycos[iw] = polyco
    peaks = sorted(peaks, key=lambda pk: pk[6], reverse=True)
    return peaks, polycos
