"""Batched search (persistent work-queue kernel, hipGraph replay) against the
CPU oracle's ffa_search on every series of a small heterogeneous batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _series(rng, n, tsamp):
    t = np.arange(n) * tsamp
    period = np.exp(rng.uniform(np.log(0.2), np.log(1.5)))
    ph = (t / period) % 1.0
    sig = np.exp(-0.5 * ((ph - 0.5) / 0.01) ** 2)
    sig *= rng.uniform(8, 30) / np.sqrt((sig ** 2).sum())
    red = np.cumsum(rng.standard_normal(n)) * 0.01
    return (rng.standard_normal(n) + red + sig).astype(np.float32)


def test_batch_matches_oracle(oracle):
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(2024)
    tsamp = 128e-6
    sizes = [int(2 ** rng.uniform(16, 18)) // 1024 * 1024 for _ in range(10)]
    bmin = np.array([240 if i % 2 else 480 for i in range(10)])
    bmax = np.where(bmin == 240, 260, 520)
    data = [_series(rng, n, tsamp) for n in sizes]
    b = BatchSearch(sizes, tsamp, bmin, bmax, 0.08, 2.0, rmed_width=1.0, group_size=4)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, x in enumerate(data):
            _, w, P, F, S = oracle.ffa_search(x, tsamp, period_min=0.08, period_max=2.0,
                                              bins_min=int(bmin[i]), bins_max=int(bmax[i]),
                                              rmed_width=1.0)
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    b.close()


def test_batch_large_downsampling_factors(oracle):
    """Periods long against bins x tsamp: the steps with downsampling factors
    above 4000 (here 6-12 s at 32-36 bins and 64 us: factors 2930-5860) take
    the direct downsampling kernel, the others the packed one; every series
    equals the oracle, eager and from the graph."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(4000)
    tsamp, pmin, pmax = 64e-6, 6.0, 12.0
    sizes = [2 ** 21, 2 ** 21 + 4096, 3 * 2 ** 20]
    bmin = np.array([32, 40, 32])
    bmax = np.array([36, 44, 36])
    data = [_series(rng, n, tsamp) for n in sizes]
    b = BatchSearch(sizes, tsamp, bmin, bmax, pmin, pmax, rmed_width=20.0, group_size=2)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    want = [oracle.ffa_search(x, tsamp, period_min=pmin, period_max=pmax, bins_min=int(bmin[i]),
                              bins_max=int(bmax[i]), rmed_width=20.0)
            for i, x in enumerate(data)]
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, (_, w, P, F, S) in enumerate(want):
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            assert (P / (F * tsamp)).max() > 4000  # a direct step is present
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    b.close()


def test_batch_many_downsampling_steps(oracle):
    """More downsampling steps than one packed launch holds (256): bins
    240-241 over P 0.1-10 s take 555 steps, in three launches; the series
    of a group with fewer steps (bins 240-260, 57 steps) take one. Every
    series equals the oracle, eager and from the graph."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(555)
    tsamp, pmin, pmax = 64e-6, 0.1, 10.0
    sizes = [2 ** 20, 2 ** 20 + 8192, 2 ** 20]
    bmin = np.array([240, 240, 240])
    bmax = np.array([241, 260, 241])
    data = [_series(rng, n, tsamp) for n in sizes]
    b = BatchSearch(sizes, tsamp, bmin, bmax, pmin, pmax, rmed_width=20.0, group_size=3)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    want = [oracle.ffa_search(x, tsamp, period_min=pmin, period_max=pmax, bins_min=int(bmin[i]),
                              bins_max=int(bmax[i]), rmed_width=20.0)
            for i, x in enumerate(data)]
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, (_, w, P, F, S) in enumerate(want):
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    b.close()


def test_batch_wide_bins_classes(oracle):
    """The reference's third search range (bins 960-1040, pipeline/config/
    example.yaml:153-154) and a wider class (2000-2200 bins) in one batch,
    with the untuned default tiling for those classes: every series equals
    the oracle, eager and from the graph."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(960)
    tsamp, pmin, pmax = 256e-6, 0.6, 5.0
    sizes = [2 ** 19, 2 ** 19 - 3333, 2 ** 19 + 2048]
    bmin = np.array([960, 2000, 960])
    bmax = np.array([1040, 2200, 1040])
    data = [_series(rng, n, tsamp) for n in sizes]
    b = BatchSearch(sizes, tsamp, bmin, bmax, pmin, pmax, rmed_width=4.0, group_size=3)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    want = [oracle.ffa_search(x, tsamp, period_min=pmin, period_max=pmax, bins_min=int(bmin[i]),
                              bins_max=int(bmax[i]), rmed_width=4.0)
            for i, x in enumerate(data)]
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, (_, w, P, F, S) in enumerate(want):
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    b.close()


def test_ffa_search_batch_api(oracle):
    import riptide_amd as rt
    from riptide_amd.batch import ffa_search_batch

    rng = np.random.default_rng(7)
    tss = [rt.TimeSeries(_series(rng, 2 ** 16 + 1000 * k, 256e-6), 256e-6) for k in range(3)]
    pgs = ffa_search_batch(tss, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                           rmed_width=1.0)
    for ts, pg in zip(tss, pgs):
        _, pg1 = rt.ffa_search(ts, period_min=0.5, period_max=2.0, bins_min=240, bins_max=260,
                               rmed_width=1.0)
        assert np.array_equal(pg.periods, pg1.periods)
        np.testing.assert_allclose(pg.snrs, pg1.snrs, rtol=1e-5, atol=1e-4)


def test_batch_tiling_classes_match_oracle(oracle):
    """Transforms of different (rows, bins) classes tiled with different
    workgroup shapes and cone depths in one batch (several runs per kernel
    variant): every series still matches the oracle."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(77)
    tsamp = 128e-6
    sizes = [2 ** 18, 2 ** 18 + 5000, 2 ** 17 + 123, 2 ** 18 - 777]
    bmin = np.array([240, 480, 240, 480])
    bmax = np.where(bmin == 240, 260, 520)
    data = [_series(rng, n, tsamp) for n in sizes]
    tiling = [[1, 60, 240, 261, 1, 0, 2],
              [60, 300, 240, 261, 2, 40, 3],
              [300, 1 << 30, 240, 261, 0, 0, 2],
              [1, 200, 480, 521, 2, 0, 4],
              [200, 1 << 30, 480, 521, 1, 0, 2]]
    b = BatchSearch(sizes, tsamp, bmin, bmax, 0.08, 2.0, rmed_width=1.0, group_size=4,
                    tiling=tiling)
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    for use_graph in (False, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, x in enumerate(data):
            _, w, P, F, S = oracle.ffa_search(x, tsamp, period_min=0.08, period_max=2.0,
                                              bins_min=int(bmin[i]), bins_max=int(bmax[i]),
                                              rmed_width=1.0)
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    b.close()


def test_batch_shared_plans_use_tile_records(oracle):
    """Series of equal length and bins range share one plan: their transforms
    take the tile records (cones and row tables built once per run) and a
    shared task list.  Every series matches the oracle, and a series searched
    in a batch of its own (no record: the tasks build their tables
    themselves, same tiling) gives bit-identical S/N."""
    import torch
    from riptide_amd import _lib
    from riptide_amd.batch import BatchSearch

    rng = np.random.default_rng(99)
    tsamp = 128e-6
    sizes = [2 ** 18, 2 ** 18, 2 ** 17 + 321, 2 ** 18, 2 ** 17 + 321, 200_000]
    bmin = np.array([480, 480, 240, 480, 240, 240])
    bmax = np.where(bmin == 240, 260, 520)
    data = [_series(rng, n, tsamp) for n in sizes]
    kw = dict(rmed_width=1.0)
    b = BatchSearch(sizes, tsamp, bmin, bmax, 0.08, 4.0, group_size=2, **kw)
    assert b.stats()["tile_records"] > 0
    xs = [_lib.to_device(x) for x in data]
    outs = b.allocate_outputs()
    b.bind(xs, outs)
    for use_graph in (False, True, True):
        for o in outs:
            o.fill_(np.nan)
        b.run(use_graph=use_graph)
        b.check()
        torch.cuda.synchronize()
        for i, x in enumerate(data):
            _, w, P, F, S = oracle.ffa_search(x, tsamp, period_min=0.08, period_max=4.0,
                                              bins_min=int(bmin[i]), bins_max=int(bmax[i]), **kw)
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
    got = [_lib.to_numpy(o) for o in outs]
    b.close()
    for i in (0, 2):
        one = BatchSearch([sizes[i]], tsamp, bmin[i:i + 1], bmax[i:i + 1], 0.08, 4.0, **kw)
        assert one.stats()["tile_records"] == 0
        o1 = one.allocate_outputs()
        one.bind([xs[i]], o1)
        one.run(use_graph=True)
        one.check()
        torch.cuda.synchronize()
        assert np.array_equal(_lib.to_numpy(o1[0]), got[i])
        one.close()


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_batch_shared_plans_every_shape(oracle, shape):
    """Shared plans (tile records, shared task lists) under every workgroup
    shape, with series that share a plan but not their trial widths (the
    widest width fixes the S/N tiling of the first series' task list; the
    others fall back where their rows do not fit), against the oracle."""
    import torch
    from riptide_amd import _lib, tune
    from riptide_amd.batch import BatchSearch
    from riptide_amd.ffautils import generate_width_trials

    rng = np.random.default_rng(1000 + shape)
    tsamp = 128e-6
    sizes = [200_000] * 4 + [150_000] * 2
    bmin = np.array([480, 480, 480, 480, 240, 240])
    bmax = np.where(bmin == 240, 260, 520)
    widths = [generate_width_trials(int(b)) for b in bmin]
    widths[1] = np.array([1, 2, 5, 11, 30], np.uint64)   # same plan as 0, other widths
    widths[3] = np.array([3, 7, 40, 80], np.uint64)
    data = [_series(rng, n, tsamp) for n in sizes]
    try:
        b = BatchSearch(sizes, tsamp, bmin, bmax, 0.08, 3.0, widths=widths, group_size=3,
                        shape=shape, rmed_width=1.0)
        assert b.stats()["tile_records"] > 0
        xs = [_lib.to_device(x) for x in data]
        outs = b.allocate_outputs()
        b.bind(xs, outs)
        b.run(use_graph=True)
        b.check()
        torch.cuda.synchronize()
        from oracle import oracle as O
        for i, x in enumerate(data):
            xn = O.normalise(O.deredden(np.asarray(x, np.float32), tsamp, 1.0, 101))
            P, F, S = O.oracle().periodogram(xn, tsamp, np.asarray(widths[i], np.uint64), 0.08,
                                             3.0, int(bmin[i]), int(bmax[i]))
            Pb, Fb = b.grid(i)
            assert np.array_equal(Pb, P) and np.array_equal(Fb, F)
            np.testing.assert_allclose(_lib.to_numpy(outs[i]), S, rtol=1e-5, atol=1e-4)
        b.close()
    finally:
        tune.set_shape(None)
