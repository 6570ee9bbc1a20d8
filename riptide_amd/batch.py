"""Batched DM-trial search on one GPU.

The reference searches DM trials one process each (pipeline/worker_pool.py:35-45
calling ffa_search, search.py:10-93).  Here a batch of series, of uneven
lengths and per-series bins ranges, is searched by a few launches per group of
series (see include/riptide_amd.h, rt_batch_*), optionally replayed from a
captured hipGraph.  Results are the same as ffa_search on each series.
"""
import ctypes

import numpy as np

from . import _lib, tune
from .ffautils import generate_width_trials
from .periodogram import Periodogram


class BatchSearch:
    """A planned batch search.  Device tensors for inputs and S/N outputs are
    bound once; ``run()`` then only enqueues work on the current stream."""

    def __init__(self, sizes, tsamps, bins_min, bins_max, period_min=1.0, period_max=30.0,
                 widths=None, ducy_max=0.20, wtsp=1.5, deredden=True, rmed_width=4.0,
                 rmed_minpts=101, already_normalised=False, group_size=64, shape=None,
                 tiling=None):
        n = len(sizes)
        self.n = n
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self.tsamps = np.ascontiguousarray(np.broadcast_to(tsamps, (n,)), dtype=np.float64)
        self.bins_min = np.ascontiguousarray(np.broadcast_to(bins_min, (n,)), dtype=np.int64)
        self.bins_max = np.ascontiguousarray(np.broadcast_to(bins_max, (n,)), dtype=np.int64)
        if widths is None:
            widths = [generate_width_trials(int(b), ducy_max=ducy_max, wtsp=wtsp)
                      for b in self.bins_min]
        self.widths = [np.asarray(w, dtype=np.uint64) for w in widths]
        flat = np.ascontiguousarray(np.concatenate(self.widths), dtype=np.uint64)
        counts = np.ascontiguousarray([w.size for w in self.widths], dtype=np.int64)
        self.period_min = float(period_min)
        self.period_max = float(period_max)
        h = ctypes.c_void_p()
        # FFA tiling: an explicit shape for every transform, else explicit
        # tiling entries [rows_lo, rows_hi, bins_lo, bins_hi, shape, cap,
        # lmax], else the autotuner's per-(rows, bins)-class table
        # (riptide_amd.tune), else the library default
        self.shape = shape
        if shape is not None:
            tiling = []
        elif tiling is None:
            tiling = tune.batch_tiling()
        self.tiling = [list(map(int, e)) for e in tiling]
        if shape is not None:
            _lib.call("rt_set_ffa_shape", int(shape))
        tune.set_batch_tiling(self.tiling)
        try:
            _lib.call("rt_batch_create", n, _lib.host_ptr(self.sizes), _lib.host_ptr(self.tsamps),
                      _lib.host_ptr(flat), _lib.host_ptr(counts), self.period_min,
                      self.period_max, _lib.host_ptr(self.bins_min), _lib.host_ptr(self.bins_max),
                      int(deredden), float(rmed_width), int(rmed_minpts),
                      int(not already_normalised), int(group_size), ctypes.byref(h))
        finally:
            tune.set_batch_tiling([])
            if shape is not None:
                tune.restore()
        self.handle = h
        self.lengths = []
        for i in range(n):
            L = ctypes.c_int64()
            _lib.call("rt_batch_length", h, i, ctypes.byref(L))
            self.lengths.append(L.value)
        self._grids = {}
        self._bound = None

    def grid(self, i):
        """(periods float64, foldbins uint32) of series i (shared by series
        with identical search parameters)."""
        key = (int(self.sizes[i]), float(self.tsamps[i]), int(self.bins_min[i]),
               int(self.bins_max[i]))
        if key not in self._grids:
            L = self.lengths[i]
            P = np.empty(L, np.float64)
            F = np.empty(L, np.uint32)
            _lib.call("rt_batch_grid", self.handle, i, _lib.host_ptr(P), _lib.host_ptr(F))
            self._grids[key] = (P, F)
        return self._grids[key]

    def allocate_outputs(self):
        return [_lib.empty((self.lengths[i], self.widths[i].size)) for i in range(self.n)]

    def bind(self, inputs, outputs):
        """inputs: n float32 device tensors (series); outputs: n float32 device
        tensors of shape (length_i, nw_i)."""
        if len(inputs) != self.n or len(outputs) != self.n:
            raise ValueError("need one input and one output per series")
        for i, (x, y) in enumerate(zip(inputs, outputs)):
            if x.numel() != self.sizes[i] or not x.is_contiguous():
                raise ValueError(f"input {i} must be a contiguous tensor of {self.sizes[i]} samples")
            if tuple(y.shape) != (self.lengths[i], self.widths[i].size) or not y.is_contiguous():
                raise ValueError(f"output {i} must have shape {(self.lengths[i], self.widths[i].size)}")
        ip = (ctypes.c_void_p * self.n)(*[x.data_ptr() for x in inputs])
        op = (ctypes.c_void_p * self.n)(*[y.data_ptr() for y in outputs])
        _lib.call("rt_batch_bind", self.handle, ip, op)
        self._bound = (inputs, outputs)

    def run(self, use_graph=True):
        _lib.call("rt_batch_run", self.handle, int(use_graph), _lib.stream())

    def run_timed(self):
        """Eager run with per-phase HIP events: dict of milliseconds."""
        ms = np.zeros(4, np.float64)
        _lib.call("rt_batch_run_timed", self.handle, _lib.stream(), _lib.host_ptr(ms))
        return {"prep_ms": float(ms[0]), "ds_ms": float(ms[1]), "ffa_ms": float(ms[2]),
                "ffa_launches": int(ms[3])}

    def check(self):
        _lib.call("rt_batch_check", self.handle, _lib.stream())

    def stats(self):
        s = np.zeros(10, np.int64)
        _lib.call("rt_batch_stats", self.handle, _lib.host_ptr(s))
        keys = ["transforms", "groups", "slots", "scratch_bytes", "ds_bytes", "fold_floats",
                "periods", "ffa_adds", "tile_records", "tile_record_bytes"]
        return dict(zip(keys, (int(v) for v in s)))

    def close(self):
        if self.handle is not None:
            _lib.call("rt_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ffa_search_batch(tseries_list, period_min=1.0, period_max=30.0, fpmin=8, bins_min=240,
                     bins_max=260, ducy_max=0.20, wtsp=1.5, deredden=True, rmed_width=4.0,
                     rmed_minpts=101, already_normalised=False, use_graph=False):
    """ffa_search over many TimeSeries in one batch; returns their Periodograms.
    `bins_min`/`bins_max` may be scalars or one value per series."""
    n = len(tseries_list)
    bmin = np.broadcast_to(np.asarray(bins_min, dtype=np.int64), (n,))
    bmax = np.broadcast_to(np.asarray(bins_max, dtype=np.int64), (n,))
    b = BatchSearch([ts.nsamp for ts in tseries_list], [ts.tsamp for ts in tseries_list], bmin,
                    bmax, period_min, period_max, ducy_max=ducy_max, wtsp=wtsp,
                    deredden=deredden, rmed_width=rmed_width, rmed_minpts=rmed_minpts,
                    already_normalised=already_normalised)
    try:
        inputs = [ts.data_device if ts.data_device is not None else _lib.to_device(ts.data)
                  for ts in tseries_list]
        outputs = b.allocate_outputs()
        b.bind(inputs, outputs)
        b.run(use_graph=use_graph)
        b.check()
        out = []
        for i, ts in enumerate(tseries_list):
            P, F = b.grid(i)
            # S/N left in HBM, copied to the host when first read
            out.append(Periodogram(b.widths[i].astype(np.int64), P.copy(), F.copy(),
                                   outputs[i], metadata=ts.metadata))
        return out
    finally:
        b.close()
