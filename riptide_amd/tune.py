"""Autotuner of the FFA kernels' tiling.

The FFA task kernels come in a few workgroup shapes (threads per workgroup,
LDS per workgroup, merge-level values each lane keeps in registers; see
``kFfaShape`` in csrc/tasks.h).  The shape fixes how many rows a workgroup
tiles (so how many passes a transform needs) and how many workgroups share a
compute unit; a transform taller than a tile is computed in passes of cone
tiles, each at most ``lmax`` merge levels deep.  Which choice is fastest
depends on the bins range, the transform's rows and the schedule (one series
or a batch), so it is measured:

    python -m riptide_amd.tune [--nsamp N] [--tsamp T] [--trials K] [--out FILE]

times, on synthetic white noise of the given shape,
 (a) the single-series periodogram under every shape, per bins class;
 (b) a batch of K series under every (shape, lmax), per (rows class, bins
     class), then the best of those under every tile cap of CAP_CHOICES
     (rows per workgroup, below what fits): the period range is cut so that
     the batch's transforms have rows in the class (rows ~ nsamp * tsamp /
     period);
and writes the winners to a JSON file (default: ``tuning.json`` next to this
module, or ``$RIPTIDE_AMD_TUNING``).  ``shape_for("periodogram", bins_min)``
and ``batch_tiling()`` read it; the periodogram entry points and
``BatchSearch`` apply it unless a shape was forced with ``set_shape``.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _lib

NUM_SHAPES = 3
DEFAULT_CLASSES = ((240, 260), (480, 520))
# rows classes of the batch tiling: [lo, hi)
ROWS_CLASSES = ((1, 128), (128, 1024), (1024, 1 << 30))
# merge levels per cone pass tried by the batch tuner (the planner allows 7)
LMAX_CHOICES = (2, 3, 4, 5)
# tile caps (rows per workgroup; 0: as many as fit) tried with the best
# (shape, lmax) of a class; a cap above what fits is clamped to it
CAP_CHOICES = (0, 96, 80, 72, 64, 56, 48)
_HERE = os.path.dirname(os.path.abspath(__file__))

_cache = None
_forced = None


def tuning_path():
    return os.environ.get("RIPTIDE_AMD_TUNING", os.path.join(_HERE, "tuning.json"))


def set_shape(shape):
    """Forces workgroup shape `shape` (0 .. NUM_SHAPES-1) for every later
    call; None returns to the tuned (or default) choice."""
    global _forced
    if shape is not None and not 0 <= int(shape) < NUM_SHAPES:
        raise ValueError(f"unknown FFA workgroup shape {shape}")
    _forced = None if shape is None else int(shape)
    _lib.call("rt_set_ffa_shape", -1 if shape is None else int(shape))


def restore():
    """Returns the library to the forced shape, or to its default."""
    _lib.call("rt_set_ffa_shape", -1 if _forced is None else _forced)


def load(path=None):
    """The tuning table (empty when none was written)."""
    global _cache
    path = path or tuning_path()
    if _cache is None or _cache[0] != path:
        table = {}
        if os.path.exists(path):
            with open(path) as f:
                table = json.load(f)
        _cache = (path, table)
    return _cache[1]


def shape_for(kind, bins_min):
    """Tuned shape for schedule `kind` ("periodogram") and the bins class
    holding bins_min (the class with the nearest lower bound), or None."""
    if _forced is not None:
        return _forced
    best = load().get(kind, {})
    if not best:
        return None
    key = min(best, key=lambda k: abs(int(k.split("-")[0]) - int(bins_min)))
    return int(best[key])


def apply(kind, bins_min):
    """Selects the tuned shape for the next launches (no-op without a table)."""
    s = shape_for(kind, bins_min)
    if s is not None and _forced is None:
        _lib.call("rt_set_ffa_shape", s)
    return s


def batch_tiling():
    """Tiling entries [rows_lo, rows_hi, bins_lo, bins_hi, shape, cap, lmax]
    of the tuning table for BatchSearch (empty without a table, or when a
    shape is forced).  A bins class [lo, hi] covers bins lo .. hi inclusive."""
    if _forced is not None:
        return []
    out = []
    for e in load().get("batch_tiles", []):
        out.append([int(e["rows"][0]), int(e["rows"][1]), int(e["bins"][0]),
                    int(e["bins"][1]) + 1, int(e["shape"]), int(e.get("cap", 0)),
                    int(e.get("lmax", 0))])
    return out


def set_batch_tiling(entries):
    """Hands tiling entries to the library for the next rt_batch_create."""
    arr = np.ascontiguousarray(np.asarray(entries, dtype=np.int32).reshape(-1, 7))
    _lib.call("rt_batch_set_tiling", _lib.host_ptr(arr), int(arr.shape[0]))


def _time_periodogram(shape, nsamp, tsamp, bins, pmin, pmax, reps=5):
    """Median wall time of one synchronous one-series periodogram call."""
    import torch
    from . import libcpp
    from .ffautils import generate_width_trials
    set_shape(shape)
    x = torch.randn(nsamp, device=_lib.device())
    w = generate_width_trials(bins[0])
    for _ in range(2):
        libcpp.periodogram_device(x, nsamp, tsamp, w, pmin, pmax, bins[0], bins[1])
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        libcpp.periodogram_device(x, nsamp, tsamp, w, pmin, pmax, bins[0], bins[1])
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def _time_batch(shape, lmax, nsamp, tsamp, bins, pmin, pmax, trials, cap=0):
    """Least of two FFA phase times of a batch of `trials` series, every
    transform tiled with (shape, lmax, cap)."""
    import torch
    from .batch import BatchSearch
    xs = [torch.randn(nsamp, device=_lib.device()) for _ in range(trials)]
    b = BatchSearch([nsamp] * trials, tsamp, bins[0], bins[1], pmin, pmax, deredden=False,
                    already_normalised=True, group_size=trials,
                    tiling=[[0, 1 << 30, 1, 1 << 30, shape, cap, lmax]])
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    b.run_timed()
    ms = min(b.run_timed()["ffa_ms"] for _ in range(2))
    b.close()
    return ms


def rows_period_range(rows, nsamp, tsamp, pmin, pmax):
    """Period range whose transforms have about `rows` = [lo, hi) rows
    (rows ~ nsamp * tsamp / period), within [pmin, pmax]; None if empty."""
    span = nsamp * tsamp
    p0 = max(pmin, span / rows[1])
    p1 = min(pmax, span / rows[0])
    return (p0, p1) if p1 > p0 * 1.001 else None


def tune(nsamp=2 ** 22, tsamp=64e-6, period_min=0.1, period_max=10.0, trials=16,
         classes=DEFAULT_CLASSES, out=None, log=print):
    """Measures every shape on every bins class and writes the table."""
    table = {"device": None, "workload": {"nsamp": nsamp, "tsamp": tsamp,
                                          "period_min": period_min, "period_max": period_max,
                                          "trials": trials},
             "periodogram": {}, "timings_ms": {}}
    try:
        import torch
        table["device"] = torch.cuda.get_device_name(0)
    except Exception:
        pass
    # every class's one-series timings first: the batch runs allocate and
    # release large scratch areas, which disturbed the one-series timings
    # taken after them
    pgs = {}
    try:
        for bins in classes:
            pgs[bins] = [_time_periodogram(s, nsamp, tsamp, bins, period_min, period_max)
                         for s in range(NUM_SHAPES)]
    finally:
        set_shape(None)
    table["batch_tiles"] = []
    for bins in classes:
        key = f"{bins[0]}-{bins[1]}"
        pg = pgs[bins]
        table["periodogram"][key] = int(np.argmin(pg))
        table["timings_ms"][key] = {"periodogram": [round(v, 3) for v in pg]}
        log(f"bins {key}: periodogram ms by shape {table['timings_ms'][key]['periodogram']}")
        for rows in ROWS_CLASSES:
            pr = rows_period_range(rows, nsamp, tsamp, period_min, period_max)
            if pr is None:
                continue
            ms = {}
            for shape in range(NUM_SHAPES):
                for lmax in LMAX_CHOICES:
                    ms[(shape, lmax, 0)] = _time_batch(shape, lmax, nsamp, tsamp, bins, pr[0],
                                                       pr[1], trials)
            # the tile cap, with the best (shape, lmax)
            s0, l0, _ = min(ms, key=ms.get)
            for cap in CAP_CHOICES:
                if (s0, l0, cap) not in ms:
                    ms[(s0, l0, cap)] = _time_batch(s0, l0, nsamp, tsamp, bins, pr[0], pr[1],
                                                    trials, cap)
            best = min(ms, key=ms.get)
            # a cap below what fits only where it is more than 1% faster than
            # the full tile (closer timings are within the run-to-run spread)
            if best[2] and ms[best] > 0.99 * ms[(s0, l0, 0)]:
                best = (s0, l0, 0)
            table["batch_tiles"].append({
                "rows": [rows[0], rows[1]], "bins": [bins[0], bins[1]], "shape": best[0],
                "cap": best[2], "lmax": best[1], "periods": [round(pr[0], 6), round(pr[1], 6)],
                "ffa_ms": {(f"shape{k[0]}_lmax{k[1]}" + (f"_cap{k[2]}" if k[2] else "")):
                           round(v, 3) for k, v in ms.items()}})
            log(f"bins {key}, rows {rows[0]}-{rows[1]} (P {pr[0]:.3g}-{pr[1]:.3g} s): best "
                f"shape {best[0]}, lmax {best[1]}, cap {best[2]}: {ms[best]:.2f} ms; shape 0, "
                f"lmax 5: {ms[(0, 5, 0)]:.2f} ms")
    set_shape(None)
    path = out or tuning_path()
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
    global _cache
    _cache = None
    log(f"wrote {path}")
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nsamp", type=int, default=2 ** 22)
    ap.add_argument("--tsamp", type=float, default=64e-6)
    ap.add_argument("--period-min", type=float, default=0.1)
    ap.add_argument("--period-max", type=float, default=10.0)
    ap.add_argument("--trials", type=int, default=16)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    tune(a.nsamp, a.tsamp, a.period_min, a.period_max, a.trials, out=a.out)


if __name__ == "__main__":
    # run the package's copy of this module: the shape forced by set_shape()
    # lives in module state that libcpp reads through `riptide_amd.tune`, which
    # is a different module object from `__main__` under `python -m`
    from riptide_amd import tune as _package_tune
    _package_tune.main()
