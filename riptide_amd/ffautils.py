"""Boxcar width trials (ffautils.py:4-11 of the reference)."""
import numpy as np


def generate_width_trials(nbins, ducy_max=0.20, wtsp=1.5):
    """Width trials 1, 2, 3, 4, 6, 9, ... : w <- max(w + 1, int(wtsp * w)) while
    w <= int(max(1, ducy_max * nbins)) (ffautils.py:4-11)."""
    trials = []
    wmax = int(max(1, ducy_max * nbins))
    w = 1
    while w <= wmax:
        trials.append(w)
        w = int(max(w + 1, wtsp * w))
    return np.asarray(trials)
