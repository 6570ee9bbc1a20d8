"""Kernel-level Python API (libffa.py:15-243 of the reference) on the GPU path."""
import numpy as np

from . import libcpp


def generate_signal(nsamp, period, phi0=0.5, ducy=0.02, amplitude=10.0, stdnoise=1.0):
    """Periodic von Mises pulse train with unit L2 norm scaled by `amplitude`,
    plus N(0, stdnoise) noise from numpy's global RNG (libffa.py:15-68; the
    same formulas and draw order, so seeded data match the reference's)."""
    kappa = np.log(2.0) / (2.0 * np.sin(np.pi * ducy / 2.0) ** 2)
    phase = (np.arange(nsamp, dtype=float) / period - phi0) * (2 * np.pi)
    signal = np.exp(kappa * (np.cos(phase) - 1.0))
    signal *= amplitude * (signal ** 2).sum() ** -0.5
    noise = np.random.normal(size=nsamp, loc=0.0, scale=stdnoise) if stdnoise > 0.0 else 0.0
    return signal + noise


def ffa2(data):
    """FFA transform of a 2-D (m periods x p bins) array, float32 (libffa.py:71-91)."""
    return libcpp.ffa2(data)


def ffa1(data, p):
    """FFA transform of a 1-D series at base period p samples; the last
    N % p samples are ignored (libffa.py:94-126)."""
    if not data.ndim == 1:
        raise ValueError("input data must be one-dimensional")
    # numpy integers are accepted too (the reference's isinstance(p, int)
    # rejects them, libffa.py:121)
    if isinstance(p, (bool, np.bool_)) or not (isinstance(p, (int, np.integer)) and p > 0):
        raise ValueError("p must be an integer > 1")
    p = int(p)
    if p > data.size:
        raise ValueError("p must be smaller than the total number of samples")
    m = data.size // p
    return ffa2(data[: m * p].reshape(m, p))


def ffafreq(N, p, dt=1.0):
    """Trial frequencies of the rows of an FFA transform (libffa.py:129-169)."""
    if not (isinstance(N, int) and N > 0):
        raise ValueError("N must be a strictly positive integer")
    if not (isinstance(p, int) and p > 1):
        raise ValueError("p must be an integer > 1")
    if not N >= p:
        raise ValueError("p must be smaller than (or equal to) N")
    if not dt > 0:
        raise ValueError("dt must be strictly positive")
    f0 = 1.0 / p
    m = N // p
    if m == 1:
        f = np.asarray([f0])
    else:
        s = np.arange(m)
        f = f0 - s / (m - 1.0) * f0 ** 2
    f /= dt
    return f


def ffaprd(N, p, dt=1.0):
    """Trial periods of the rows of an FFA transform (libffa.py:172-191)."""
    return 1.0 / ffafreq(N, p, dt=dt)


def boxcar_snr(data, widths, stdnoise=1.0):
    """S/N of profile(s) along the last axis for each boxcar width; output
    shape data.shape[:-1] + (len(widths),) (libffa.py:194-225)."""
    widths = np.asarray(widths, dtype=np.uint64)
    b = data.shape[-1]
    x = data.reshape(-1, b).astype(np.float32)
    snr = libcpp.snr2(x, widths, stdnoise)
    return snr.reshape(list(data.shape[:-1]) + [widths.size])


def downsample(data, factor):
    """Downsample by a real-valued factor (libffa.py:228-243)."""
    return libcpp.downsample(data, factor)
