"""DM-trial search driver on GPU batches: the reference's WorkerPool
(pipeline/worker_pool.py:10-68) loads, dereddens, normalises and searches one
DM trial per process and finds peaks in each periodogram; here the trials
are staged to the device in chunks (staging.stage_files), searched as one
BatchSearch per chunk and search range, and their peaks found on the device
(peak_detection.PeakFinder), with the same configuration dictionaries."""
import numpy as np

from .batch import BatchSearch
from .peak_detection import PeakFinder
from .staging import stage_files


def _chunks(seq, n):
    for i in range(0, len(seq), n):
        yield seq[i:i + n]


class BatchWorker:
    """deredden_params: {"rmed_width", "rmed_minpts"}; range_confs: list of
    {"ffa_search": {...}, "find_peaks": {...}} as in the reference's YAML
    configuration; batch_size: DM trials per GPU batch.

    The batch of each search range (its plans, task lists, tile records and
    device buffers) and its peak finders are kept and rebound while the next
    chunk holds series of the same lengths and sampling times, as the DM
    trials of one observation do; close() releases them."""

    def __init__(self, deredden_params, range_confs, batch_size=64, fmt="presto"):
        if fmt not in ("presto", "sigproc"):
            raise ValueError(f"unknown time series format {fmt!r}")
        self.deredden_params = dict(deredden_params)
        self.range_confs = list(range_confs)
        self.batch_size = int(batch_size)
        self.fmt = fmt
        self._kept = {}  # range index -> (series key, BatchSearch, S/N outputs, peak finders)

    def close(self):
        """Releases the kept batches."""
        for _, b, _, _ in self._kept.values():
            b.close()
        self._kept.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _batch(self, ci, kw, staged, nfiles):
        """The batch of range `ci` for these series: the kept one when the
        series lengths and sampling times are those it was made for."""
        key = (tuple(int(n) for n in staged.sizes), tuple(float(t) for t in staged.tsamps))
        kept = self._kept.get(ci)
        if kept is not None and kept[0] == key:
            return kept[1], kept[2], kept[3]
        if kept is not None:
            del self._kept[ci]
            kept[1].close()
        rw = float(self.deredden_params["rmed_width"])
        rm = int(self.deredden_params.get("rmed_minpts", 101))
        b = BatchSearch(staged.sizes, staged.tsamps, kw.get("bins_min", 240),
                        kw.get("bins_max", 260), kw.get("period_min", 1.0),
                        kw.get("period_max", 30.0), ducy_max=kw.get("ducy_max", 0.20),
                        wtsp=kw.get("wtsp", 1.5), deredden=True, rmed_width=rw,
                        rmed_minpts=rm, already_normalised=False,
                        group_size=min(64, nfiles))
        try:
            snrs = b.allocate_outputs()
        except Exception:
            b.close()
            raise
        self._kept[ci] = (key, b, snrs, {})
        return b, snrs, self._kept[ci][3]

    def process_fname_list(self, fnames):
        """All peaks of all files, as WorkerPool.process_fname_list."""
        peaks = []
        for chunk in _chunks(list(fnames), self.batch_size):
            peaks.extend(self.process_chunk(chunk))
        return peaks

    def process_chunk(self, fnames):
        staged = stage_files(fnames)
        out = []
        for ci, conf in enumerate(self.range_confs):
            kw = dict(conf["ffa_search"])
            b, snrs, finders = self._batch(ci, kw, staged, len(fnames))
            b.bind(staged.series, snrs)
            b.run(use_graph=False)
            b.check()
            fp = dict(conf.get("find_peaks", {}))
            segwidth = fp.pop("segwidth", 5.0)
            for i in range(len(fnames)):
                P, F = b.grid(i)
                tobs = staged.sizes[i] * staged.tsamps[i]
                key = (staged.sizes[i], staged.tsamps[i], segwidth)
                if key not in finders:
                    finders[key] = PeakFinder(P, F, b.widths[i].astype(np.int64), tobs,
                                              segwidth=segwidth)
                pk, _ = finders[key].find(snrs[i], dm=staged.metadata[i]["dm"], **fp)
                out.extend(pk)
        return out
