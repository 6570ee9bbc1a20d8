"""Synthetic DM-trial series of the search configurations: white N(0, 1)
noise, red noise with a power-law spectrum, and Gaussian pulse trains scaled
to a true S/N by the L2-norm convention of generate_signal (libffa.py:57-59:
the noiseless pulse train has L2 norm equal to its S/N against unit white
noise).  numpy (host) and torch (device) versions draw from different
generators; each is reproducible from its seed."""
import numpy as np


def red_noise_shape(nfft, tsamp, index, fref=0.1):
    """Amplitude shaping of an rFFT of length nfft: (f / fref)^(index/2),
    zero at DC, so that the shaped white noise has a PSD of f^index equal to
    the white level at fref Hz."""
    f = np.fft.rfftfreq(nfft, d=tsamp)
    shape = np.zeros_like(f)
    shape[1:] = (f[1:] / fref) ** (0.5 * index)
    return shape


def gaussian_pulses(nsamp, tsamp, period, duty, snr, phase0=0.5):
    """Train of Gaussian pulses (FWHM = duty * period) centred at phase
    phase0, float64, with L2 norm `snr`."""
    t = np.arange(nsamp) * tsamp
    ph = np.remainder(t / period, 1.0)
    d = ph - phase0
    d -= np.round(d)  # circular phase distance
    sig = np.exp(-0.5 * (d / (duty / 2.355)) ** 2)
    return sig * (snr / np.sqrt((sig ** 2).sum()))


def series(nsamp, tsamp, seed, red_index=-1.5, pulses=(), fref=0.1):
    """float32 series: white N(0,1) + red noise (PSD index red_index, equal
    to the white level at fref Hz; red_index None: none) + the Gaussian pulse
    trains `pulses` = [(period s, duty, snr), ...]."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(nsamp)
    if red_index is not None:
        nfft = 1 << int(np.ceil(np.log2(nsamp)))
        w = np.fft.rfft(rng.standard_normal(nfft))
        x += np.fft.irfft(w * red_noise_shape(nfft, tsamp, red_index, fref), n=nfft)[:nsamp]
    for period, duty, snr in pulses:
        x += gaussian_pulses(nsamp, tsamp, period, duty, snr)
    return x.astype(np.float32)


def series_torch(torch, nsamp, tsamp, generator, dev, red_index=-1.5, pulses=(), nfft=None,
                 fref=0.1):
    """The same model generated on the device (torch's generator): returns a
    float32 device tensor.  nfft: FFT length of the red noise (default: the
    next power of two of nsamp; a batch passes one length for all series)."""
    nfft = nfft or 1 << int(np.ceil(np.log2(nsamp)))
    x = torch.randn(nsamp, generator=generator, device=dev)
    if red_index is not None:
        shape = torch.as_tensor(red_noise_shape(nfft, tsamp, red_index, fref), device=dev,
                                dtype=torch.float32)
        w = torch.fft.rfft(torch.randn(nfft, generator=generator, device=dev))
        x += torch.fft.irfft(w * shape, n=nfft)[:nsamp]
    for period, duty, snr in pulses:
        t = torch.arange(nsamp, device=dev, dtype=torch.float64) * tsamp
        d = torch.remainder(t / period, 1.0) - 0.5
        d -= torch.round(d)
        sig = torch.exp(-0.5 * (d / (duty / 2.355)) ** 2)
        sig *= snr / torch.sqrt((sig ** 2).sum())
        x += sig.float()
    return x.contiguous()
