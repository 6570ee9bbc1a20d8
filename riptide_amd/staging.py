"""Batched input staging: many dedispersed time series files (PRESTO .inf
+ .dat, SIGPROC .tim) read into one pinned host buffer, moved to the device
in one asynchronous transfer, 8-bit samples widened to float32 there.  The
reference reads one file at a time into host float32 arrays
(time_series.py:301-377) for a pool of per-DM-trial processes; the result
here feeds BatchSearch directly."""
import os

import numpy as np

from . import _lib
from .metadata import Metadata

_ALIGN = 256


def _torch():
    import torch
    return torch


def _describe(fname):
    """(data file, byte offset, sample count, nbits, signed, tsamp, metadata)."""
    if fname.endswith(".inf"):
        from .reading.presto import PrestoInf
        inf = PrestoInf(fname)
        n = os.path.getsize(inf.data_fname) // 4
        return inf.data_fname, 0, n, 32, False, inf["tsamp"], Metadata.from_presto_inf(inf)
    from .reading.sigproc import SigprocHeader
    sh = SigprocHeader(fname)
    meta = Metadata.from_sigproc(sh)
    return (sh.fname, sh.bytesize, sh.nsamp, sh["nbits"], bool(sh.get("signed", False)),
            sh["tsamp"], meta)


class StagedBatch:
    """Device float32 series (views into one device buffer) with their
    sampling times and metadata."""

    def __init__(self, series, tsamps, metadata, fnames):
        self.series = series
        self.tsamps = tsamps
        self.metadata = metadata
        self.fnames = fnames

    @property
    def sizes(self):
        return [int(s.numel()) for s in self.series]


def stage_files(fnames, stream=None):
    """Reads every file of `fnames` and returns a StagedBatch whose series are
    on the current device (the copy and conversions are enqueued on `stream`,
    default the current stream; the host buffer is kept until they finish)."""
    torch = _torch()
    dev = _lib.device()
    descs = [_describe(f) for f in fnames]
    offs, total = [], 0
    for (_, _, n, nbits, _, _, _) in descs:
        offs.append(total)
        total += -(-n * nbits // 8 // _ALIGN) * _ALIGN if n else 0
    host = torch.empty(max(total, 1), dtype=torch.uint8, pin_memory=True)
    hv = host.numpy()
    for (path, start, n, nbits, _, _, _), o in zip(descs, offs):
        nbytes = n * nbits // 8
        with open(path, "rb") as f:
            f.seek(start)
            got = f.readinto(memoryview(hv[o:o + nbytes]))
        if got != nbytes:
            raise IOError(f"short read from {path!r}")
    s = stream or torch.cuda.current_stream()
    raw = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.stream(s):
        raw.copy_(host, non_blocking=True)
        out = []
        for (path, start, n, nbits, signed, _, _), o in zip(descs, offs):
            if nbits == 32:
                out.append(raw[o:o + 4 * n].view(torch.float32))
            else:
                y = torch.empty(n, dtype=torch.float32, device=dev)
                _lib.call("rt_convert_samples", _lib.ptr(raw[o:]), n, 8, 1 if signed else 0,
                          _lib.ptr(y), _lib.stream_handle(s))
                out.append(y)
    batch = StagedBatch(out, [d[5] for d in descs], [d[6] for d in descs], list(fnames))
    batch._keep = (host, raw)  # alive until the batch is dropped
    return batch
