"""Friends-of-friends clustering in one dimension (clustering.py:4-48 of the
reference; same interface and result: points closer than r to a neighbour in
sorted order share a cluster).  Runs on the host: it only sees the few
periods that pass the peak selection threshold."""
import numpy as np


def cluster1d(x, r, already_sorted=False):
    """List of clusters of x, each an array of indices into x."""
    if not len(x):
        return []
    if already_sorted:
        order = np.arange(len(x))
        gaps = np.diff(x)
    else:
        order = x.argsort()
        gaps = np.diff(x[order])
    cuts = np.where(abs(gaps) > r)[0]
    if not len(cuts):
        return [order]
    edges = np.concatenate(([0], cuts + 1, [len(x)]))
    return [order[a:b] for a, b in zip(edges[:-1], edges[1:])]
