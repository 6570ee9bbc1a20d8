"""Oracle of candidate folding (test infrastructure, CPU only): folding.py:
6-85 of the reference restated with the oracle's C downsampling (the
reference's downsample.hpp arithmetic) and the same numpy operations."""
import numpy as np

from .oracle import oracle


def downsample_vertical(X, factor):
    m = X.shape[0]
    if not factor > 1:
        raise ValueError("factor must be > 1")
    if not factor < m:
        raise ValueError("factor must be strictly smaller than the number of input lines")
    cols = np.ascontiguousarray(X.T)
    out = np.asarray([oracle().downsample(c, factor) for c in cols])
    return np.ascontiguousarray(out.T)


def fold(data, tsamp, period, bins, subints=None):
    length = data.size * tsamp
    if period > length:
        raise ValueError("Period exceeds data length")
    tbin = period / bins
    if not tbin > tsamp:
        raise ValueError("Bin width is shorter than sampling time")
    if subints is not None:
        subints = int(subints)
        if not subints >= 1:
            raise ValueError("subints must be >= 1 or None")
        if subints > length / period:
            raise ValueError("subints exceeds the number of signal periods")
    factor = tbin / tsamp
    down = oracle().downsample(np.ascontiguousarray(data, np.float32), factor)
    m = down.size // bins
    folded = down[:m * bins].reshape(m, bins)
    folded *= (m * factor) ** -0.5
    if subints == 1 or m == 1:
        return folded.sum(axis=0)
    if subints is None or subints == m:
        return folded
    return downsample_vertical(folded, m / subints)
