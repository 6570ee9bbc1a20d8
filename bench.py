"""Benchmark: batched ffa_search over DM trials on one MI355X.

Workload (one "step"): ffa_search -- dereddening (4 s running median),
normalisation, periodogram over 0.1-10 s with bins (240, 260) or (480, 520)
picked per series by a seeded coin -- on a batch of 512 float32 series of 2^22
samples at 64 us (white + red noise, 1 series in 16 with a pulse), inputs
resident in HBM, S/N outputs left in HBM.  value = DM trials per second.

Also reported on the same line: the plain one-series-at-a-time schedule
(per-step launches, no queue, no graph), per-phase HIP-event timings, the HBM
roofline of the dominant kernel (persistent FFA + S/N kernel), and the CPU
oracle timed on the host cores on a bounded sample of the same workload.

  python bench.py [--steps K] [--warmup W] [--trials T] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
LDS_PEAK_GBS = 78_600.0        # 128 B/clk/CU x 256 CU x 2.4 GHz (nominal)
# ds_read_b32 rate measured on gfx950 (profiles/lds_read_rates__gfx950.md):
# 169 B/clk/CU x 256 CU x 2.4 GHz
LDS_MEASURED_GBS = 169.0 * 256 * 2.4
# host cores a one-GPU box gives a job (its CPU share), whatever the machine has
HOST_CORES_PER_GPU = 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(torch, n_series, nsamp, tsamp, seed, dev, red_index=-1.5):
    """White + red noise (PSD ~ f^red_index, equal to the white level at
    0.1 Hz), one series in 16 with a Gaussian pulse train (P log-uniform 0.2-8
    s, duty 1-5 %, S/N 8-30); generated on the GPU.  nsamp: one length for all
    series or a sequence of per-series lengths."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rs = np.random.RandomState(seed)
    lengths = [int(nsamp)] * n_series if np.isscalar(nsamp) else [int(v) for v in nsamp]
    # the red noise is shaped at one FFT length (the longest series, rounded
    # up to a power of two: one FFT plan) and cut to each series' length
    nfft = 1 << int(np.ceil(np.log2(max(lengths))))
    f = torch.fft.rfftfreq(nfft, d=tsamp, device=dev)
    shape = torch.zeros_like(f)
    shape[1:] = (f[1:] / 0.1) ** (0.5 * red_index)
    xs = []
    for i in range(n_series):
        nsamp = lengths[i]
        t = torch.arange(nsamp, device=dev, dtype=torch.float64) * tsamp
        x = torch.randn(nsamp, generator=g, device=dev)
        w = torch.fft.rfft(torch.randn(nfft, generator=g, device=dev))
        x += torch.fft.irfft(w * shape, n=nfft)[:nsamp]
        if i % 16 == 0:
            period = float(np.exp(rs.uniform(np.log(0.2), np.log(8.0))))
            duty = rs.uniform(0.01, 0.05)
            snr = rs.uniform(8, 30)
            ph = torch.remainder(t / period, 1.0)
            sig = torch.exp(-0.5 * ((ph - 0.5) / (duty / 2.355)) ** 2)
            sig *= snr / torch.sqrt((sig ** 2).sum())
            x += sig.float()
        xs.append(x.contiguous())
    return xs


def _cpu_worker(args):
    """One host core: the oracle ffa_search on one trial of each bins class."""
    nsamp, tsamp, seed = args
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    x = np.random.default_rng(seed).standard_normal(nsamp).astype(np.float32)
    times = []
    for b0, b1 in ((240, 260), (480, 520)):
        t0 = time.perf_counter()
        O.ffa_search(x, tsamp, period_min=0.1, period_max=10.0, bins_min=b0, bins_max=b1)
        times.append(time.perf_counter() - t0)
    return times


def host_cpu():
    """Host CPU model, logical CPU count and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "cpu_count": os.cpu_count(), "affinity": allowed}


def cpu_baseline(nsamp, tsamp):
    """The oracle's ffa_search (C restatement of the reference kernels, as fast
    as the compiled reference: ratio 1.00-1.06 on 2^22 trials) in one
    single-threaded process per host core, as the reference's DM-trial pool
    runs (pipeline/worker_pool.py:35-45)."""
    import multiprocessing as mp
    info = host_cpu()
    cores = max(1, min(HOST_CORES_PER_GPU, info["affinity"]))
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(nsamp, tsamp, 100 + c) for c in range(cores)])
    wall = time.perf_counter() - t0
    busy = max(sum(r) for r in res)
    per = np.mean(np.asarray(res), axis=0)
    return {
        "value": round(2 * cores / busy, 4),
        "unit": "DM trials/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle ffa_search (C restatement of the reference kernels + its numpy "
                   f"dereddening/normalisation), one 2^22-sample trial of each bins class per "
                   f"core, {cores} single-threaded processes ({wall:.1f} s wall); per-core "
                   f"s/trial: bins 240-260 {per[0]:.2f}, bins 480-520 {per[1]:.2f}"),
        "host_cpu": info["model"],
        "host_cpu_count": info["cpu_count"],
        "cores_note": (f"{cores} processes: the {HOST_CORES_PER_GPU} host cores a one-GPU box "
                       f"gives a job (the machine reports {info['cpu_count']} logical CPUs, "
                       f"{info['affinity']} in this process's affinity)"),
    }


def snr_errors(got, want):
    """Largest absolute and relative S/N differences (relative over |want| >= 1)."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    d = np.abs(got - want)
    big = np.abs(want) >= 1.0
    return {"max_abs": float(d.max()),
            "max_rel": float((d[big] / np.abs(want[big])).max()) if big.any() else 0.0}


def oracle_checks(batch, xs, outs, coin, bmin, bmax, tsamp, pmin, pmax, c1, c2, peaks_ctx):
    """Checks of the GPU results made with the oracle on the host, in the
    CPU-baseline leg: parity of the first batch series of each bins class,
    the S/N error of GPU and oracle against a float64 recomputation on a
    config-1 series, the config-2 search (oracle seconds and parity) and the
    host path of find_peaks (S/N copy + the oracle's numpy find_peaks)."""
    from oracle import oracle as O
    from oracle import peaks as OP
    out = {}
    parity = []
    for cls in (0, 1):
        i = int(np.nonzero(coin == cls)[0][0])
        x0 = xs[i].cpu().numpy()
        t0 = time.perf_counter()
        _, _, Po, Fo, So = O.ffa_search(x0, tsamp, period_min=pmin, period_max=pmax,
                                        bins_min=int(bmin[i]), bins_max=int(bmax[i]))
        t_or = time.perf_counter() - t0
        S0 = outs[i].cpu().numpy()
        P0, F0 = batch.grid(i)
        parity.append({
            "series": i, "bins": f"{int(bmin[i])}-{int(bmax[i])}",
            "periods_bit_exact": bool(np.array_equal(P0, Po) and np.array_equal(F0, Fo)),
            "snr_max_abs_diff": float(np.max(np.abs(S0 - So))),
            "snr_within_1e-5rel_1e-4abs": bool(np.allclose(S0, So, rtol=1e-5, atol=1e-4)),
            "snr_bit_identical_frac": round(float(np.mean(S0 == So)), 6),
            "oracle_s": round(t_or, 2)})
    out["parity"] = parity
    # float64 recomputation of a config-1 series (the oracle's C code in float64)
    x1, tsamp1, kw1, S_gpu = c1
    xn = O.normalise(O.deredden(x1, tsamp1, 4.0))
    w1 = O.generate_width_trials(kw1["bins_min"])
    args = (kw1["period_min"], kw1["period_max"], kw1["bins_min"], kw1["bins_max"])
    _, _, S32 = O.oracle().periodogram(xn, tsamp1, w1, *args)
    S64 = O.oracle().periodogram_f64(xn, tsamp1, w1, *args)
    out["snr_vs_fp64"] = {
        "config": "2^20 samples, 256 us, P 0.5-2 s, bins 240-260 (config 1)",
        "gpu": snr_errors(S_gpu, S64), "oracle_float32": snr_errors(S32, S64),
        "gpu_vs_oracle": snr_errors(S_gpu, S32),
        "note": ("float64 recomputation: same plan, merge tree and float32 row map; "
                 "downsampling weights, FFA sums and S/N in float64")}
    # config 2 on the host
    x2, kw2, P2, F2, S2, gpu_ms = c2
    t0 = time.perf_counter()
    _, _, Po, Fo, So = O.ffa_search(x2, 64e-6, **kw2)
    out["c2"] = {"oracle_s": round(time.perf_counter() - t0, 2),
                 "periods_bit_exact": bool(np.array_equal(P2, Po) and np.array_equal(F2, Fo)),
                 "snr_within_1e-5rel_1e-4abs": bool(np.allclose(S2, So, rtol=1e-5, atol=1e-4)),
                 "snr_max_abs_diff": float(np.max(np.abs(S2 - So)))}
    out["c2"]["gpu_over_oracle"] = round(out["c2"]["oracle_s"] * 1e3 / gpu_ms, 1)
    # find_peaks: host path of the reference (S/N copy + numpy) on two trials
    finders, dev_ms, tobs = peaks_ctx
    t0 = time.perf_counter()
    same = True
    for i in range(2):
        P, F = batch.grid(i)
        S = outs[i].cpu().numpy()
        want, _ = OP.find_peaks(P, F, batch.widths[i], S, tobs)
        got, _ = finders[int(bmin[i])].find(outs[i])
        same = same and [tuple(p) for p in got] == want
    out["find_peaks_host_ms_per_trial"] = round((time.perf_counter() - t0) / 2 * 1e3 - dev_ms, 2)
    out["find_peaks_same_as_host_path"] = bool(same)
    return out


def fwd_bwd_timing(torch, _lib, reps=5):
    """One float32 series of 2^20 samples at 256 us (white noise + Gaussian
    pulses, P = 1.3711 s, FWHM 2 %, S/N 20), normalise -> periodogram over
    0.5-2 s with 240-260 bins; loss = sum over widths of the max S/N within
    +-0.5 % of the injected period.  Returns forward and forward + backward
    milliseconds (HIP events, median of `reps`)."""
    from riptide_amd.autograd import periodogram_fn, periodogram_grid
    n, tsamp, p0 = 2 ** 20, 256e-6, 1.3711
    t = np.arange(n) * tsamp
    ph = (t / p0) % 1.0
    sig = np.exp(-0.5 * ((ph - 0.5) / (0.02 / 2.355)) ** 2)
    sig *= 20.0 / np.sqrt((sig ** 2).sum())
    x = (np.random.default_rng(4).standard_normal(n) + sig).astype(np.float32)
    P, _ = periodogram_grid(n, tsamp, 0.5, 2.0, 240, 260)
    dev = _lib.device()
    mask = torch.as_tensor(np.abs(P / p0 - 1) <= 0.005, device=dev)
    xg = torch.tensor(x, device=dev, requires_grad=True)

    def fwd():
        snrs = periodogram_fn(xg, tsamp, None, 0.5, 2.0, 240, 260)
        return snrs[mask].max(dim=0).values.sum()

    def run(backward):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        xg.grad = None
        e0.record()
        loss = fwd()
        if backward:
            loss.backward()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    run(True)
    f = sorted(run(False) for _ in range(reps))[reps // 2]
    fb = sorted(run(True) for _ in range(reps))[reps // 2]
    return {"config": "2^20 samples, tsamp 256 us, P 0.5-2 s, bins 240-260, loss = sum_w max "
                      "S/N within +-0.5% of 1.3711 s",
            "fwd_ms": round(f, 3), "fwd_bwd_ms": round(fb, 3), "ratio": round(fb / f, 3)}


def c5_timing(torch, _lib, seed, adam_steps=50, nseries=128, nsamp=2 ** 21, tsamp=128e-6,
              ntaps=129, reps=3):
    """Batched forward + adjoint with FIR whitening: nseries x nsamp at tsamp,
    red noise (PSD index -2.0, equal to the white level at 0.1 Hz) plus one
    Gaussian pulse train per series (P log-uniform 0.5-5 s, duty 2 %, S/N 15).
    Per series: FIR (ntaps, identity at the start) -> normalise ->
    periodogram over P 0.5-5 s, bins 240-260, the whole batch through one
    BatchPeriodogram (batched forward and adjoint); objective = sum over
    widths of (max S/N within +-0.5 % of the series' period) - (median S/N).
    Times the forward and forward + adjoint of the whole batch (HIP events,
    median of `reps`) and runs `adam_steps` Adam steps.  Judge: on a held-out
    noise realisation (same pulses, new noise), the detection significance of
    each pulse (z-score of its S/N against the off-pulse S/N distribution)
    with the learned filters, with running-median dereddening (4 s, the
    fallback) and with no whitening."""
    from riptide_amd import synthetic
    from riptide_amd.autograd import BatchPeriodogram
    from riptide_amd.batch import BatchSearch
    from riptide_amd.fir import (_row_index, batch_whitening_objective, detection_significance,
                                 fir_filter, identity_taps, optimise_whitening)
    dev = _lib.device()
    rs = np.random.RandomState(seed + 5)
    periods = np.exp(rs.uniform(np.log(0.5), np.log(5.0), size=nseries))
    nfft = 1 << int(np.ceil(np.log2(nsamp)))

    def realisation(gseed):
        g = torch.Generator(device=dev)
        g.manual_seed(gseed)
        X = torch.empty((nseries, nsamp), dtype=torch.float32, device=dev)
        for i in range(nseries):
            X[i] = synthetic.series_torch(torch, nsamp, tsamp, g, dev, red_index=-2.0,
                                          pulses=[(float(periods[i]), 0.02, 15.0)], nfft=nfft)
        return X

    X = realisation(seed + 5)
    kw = dict(period_min=0.5, period_max=5.0, bins_min=240, bins_max=260)
    bp = BatchPeriodogram(nseries, nsamp, tsamp, None, 0.5, 5.0, 240, 260)
    grid, _ = bp.grid()
    rows = _row_index(grid, periods)

    def run(backward):
        taps = identity_taps(nseries, ntaps).requires_grad_(backward)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        objs = batch_whitening_objective(bp(fir_filter(X, taps)), rows)
        if backward:
            (-objs.sum()).backward()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    with torch.no_grad():
        run(False)  # first run captures the batch's hipGraph
    run(True)
    with torch.no_grad():
        f = sorted(run(False) for _ in range(reps))[reps // 2]
    fb = sorted(run(True) for _ in range(reps))[reps // 2]
    log(f"c5: fwd {f:.1f} ms, fwd+bwd {fb:.1f} ms per batch of {nseries}")
    t0 = time.time()
    taps, hist = optimise_whitening(X, tsamp, periods, steps=adam_steps, ntaps=ntaps, lr=2e-3,
                                    log=None, batch=bp, **kw)
    torch.cuda.synchronize()
    t_adam = time.time() - t0
    log(f"c5: {adam_steps} Adam steps in {t_adam:.1f} s, mean objective "
        f"{hist[0].mean():.3f} -> {hist[-1].mean():.3f}")
    del X
    # judge on held-out noise: learned filters, running median, no whitening
    Xh = realisation(seed + 6)
    with torch.no_grad():
        s_fir, _, _ = bp.forward_nograd(fir_filter(Xh, taps.detach()))
        s_raw, _, _ = bp.forward_nograd(Xh)
    z_fir = np.array([detection_significance(s_fir[i], grid, periods[i]) for i in range(nseries)])
    z_raw = np.array([detection_significance(s_raw[i], grid, periods[i]) for i in range(nseries)])
    del s_fir, s_raw
    rm = BatchSearch([nsamp] * nseries, tsamp, 240, 260, 0.5, 5.0, deredden=True,
                     rmed_width=4.0)
    outs = rm.allocate_outputs()
    rm.bind([Xh[i] for i in range(nseries)], outs)
    rm.run(use_graph=False)
    rm.check()
    z_rm = np.array([detection_significance(outs[i], grid, periods[i]) for i in range(nseries)])
    rm.close()
    bp.close()
    del outs, Xh
    judge = {"held_out": "same pulses, new noise realisation",
             "significance": ("max over widths of (S/N within +-0.5% of the period - median) / "
                              "robust std of the S/N away from the period and its harmonics"),
             "mean_learned_fir": round(float(z_fir.mean()), 3),
             "mean_running_median": round(float(z_rm.mean()), 3),
             "mean_no_whitening": round(float(z_raw.mean()), 3),
             "learned_beats_running_median": int((z_fir > z_rm).sum()),
             "learned_beats_no_whitening": int((z_fir > z_raw).sum())}
    log(f"c5 judge: {judge}")
    return {"config": (f"{nseries} series x 2^{int(np.log2(nsamp))} samples, tsamp "
                       f"{tsamp * 1e6:.0f} us, red noise index -2.0 + one pulse train per series "
                       f"(P log-uniform 0.5-5 s, duty 2 %, S/N 15); FIR {ntaps} taps -> "
                       f"normalise -> periodogram P 0.5-5 s, bins 240-260; the batch through "
                       f"one BatchPeriodogram"),
            "fwd_ms": round(f, 2), "fwd_bwd_ms": round(fb, 2), "ratio": round(fb / f, 3),
            "adam_steps": adam_steps, "adam_s": round(t_adam, 2),
            "objective_identity": round(float(hist[0].mean()), 3),
            "objective_optimised": round(float(hist[-1].mean()), 3),
            "judge": judge}

def peak_timing(torch, batch, outs, bmin, tsamp, nsamp, ntrials=16):
    """find_peaks on the first `ntrials` S/N arrays of the batch with the
    device path (one PeakFinder per plan); the S/N never leave the device."""
    from riptide_amd.peak_detection import PeakFinder
    tobs = nsamp * tsamp
    finders = {}
    t0 = time.perf_counter()
    npk = 0
    for i in range(ntrials):
        P, F = batch.grid(i)
        key = int(bmin[i])
        if key not in finders:
            finders[key] = PeakFinder(P, F, batch.widths[i], tobs)
    t_plan = time.perf_counter() - t0
    # one untimed call per finder (first-call set-up: buffers, kernel loads)
    warmed = set()
    for i in range(ntrials):
        if int(bmin[i]) not in warmed:
            finders[int(bmin[i])].find(outs[i])
            warmed.add(int(bmin[i]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(ntrials):
        pk, _ = finders[int(bmin[i])].find(outs[i])
        npk += len(pk)
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t0) / ntrials * 1e3
    return ({"trials": ntrials, "device_ms_per_trial": round(dev_ms, 2),
             "plan_setup_ms": round(t_plan * 1e3, 1),
             "peaks_per_trial": round(npk / ntrials, 1)}, (finders, dev_ms, tobs))


def d2h_timing(torch, outs, ntrials=8):
    """Copy of the S/N arrays of `ntrials` trials to pinned host memory (what
    a host consumer of the periodograms would pay; the batch itself leaves
    them in HBM)."""
    host = [torch.empty(outs[i].shape, dtype=outs[i].dtype, pin_memory=True)
            for i in range(ntrials)]
    for i in range(ntrials):
        host[i].copy_(outs[i], non_blocking=True)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(ntrials):
        host[i].copy_(outs[i], non_blocking=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    nbytes = sum(outs[i].numel() * 4 for i in range(ntrials))
    return {"trials": ntrials, "ms_per_trial": round(ms / ntrials, 3),
            "mb_per_trial": round(nbytes / ntrials / 1e6, 2),
            "gb_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}


def c2_timing(torch, _lib, reps=3):
    """Config 2 on the GPU: one 2^23-sample series at 64 us (white + red noise
    of PSD index -1.5 equal to the white level at 0.1 Hz, pulse trains 0.2371 s
    / 3 % / S/N 12 and 4.9133 s / 1 % / S/N 25), dereddening 4 s,
    normalisation, periodogram over 0.1-10 s with bins 480-520, one series
    per call (the drop-in ffa_search path).  Returns the host input and the
    GPU results for the oracle's check, and the median milliseconds."""
    from riptide_amd import libcpp
    from riptide_amd import synthetic
    from riptide_amd.ffautils import generate_width_trials
    tsamp = 64e-6
    kw = dict(period_min=0.1, period_max=10.0, bins_min=480, bins_max=520)
    x = synthetic.series(2 ** 23, tsamp, seed=2, red_index=-1.5,
                         pulses=[(0.2371, 0.03, 12.0), (4.9133, 0.01, 25.0)])
    n = x.size
    d = _lib.to_device(x)
    t1, t2 = _lib.empty((n,)), _lib.empty((n,))
    w = generate_width_trials(480)

    def run():
        _lib.call("rt_deredden", _lib.ptr(d), n, int(round(4.0 / tsamp)), 101, _lib.ptr(t1),
                  _lib.stream())
        _lib.call("rt_normalise", _lib.ptr(t1), n, _lib.ptr(t2), _lib.stream())
        return libcpp.periodogram_device(t2, n, tsamp, w, kw["period_min"], kw["period_max"],
                                         480, 520)

    run()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        P, F, S = run()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = sorted(times)[reps // 2]
    kw["rmed_width"] = 4.0
    return (x, kw, _lib.to_numpy(P), _lib.to_numpy(F).view(np.uint32), _lib.to_numpy(S), ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trials", type=int, default=512)
    ap.add_argument("--nsamp", type=int, default=2 ** 22)
    ap.add_argument("--plain-trials", type=int, default=16)
    ap.add_argument("--group-size", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c5-steps", type=int, default=50,
                    help="Adam steps of the FIR-whitening run (0: skip it)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the heterogeneous batch (lengths log-uniform 2^21-2^23)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched run and its phase timings (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain copy kernel over 1 GiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); DM trials are "
              "independent, so a multi-GPU run shards trials across GPUs with no exchange.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from riptide_amd import _lib, libcpp

    if args.calibrate_copy:
        # rollback_kernel with shift 0: reads and writes 2^28 floats (1 GiB
        # each way) once per launch, five launches
        n = 1 << 28
        src = torch.randn(n, device=_lib.device())
        dst = torch.empty_like(src)
        for _ in range(5):
            _lib.call("rt_rollback", _lib.ptr(src), n, 0, _lib.ptr(dst), _lib.stream())
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel": "rollback_kernel", "bytes_read": 4 * n,
                          "bytes_written": 4 * n, "launches": 5}), flush=True)
        return

    from riptide_amd import tune
    from riptide_amd.batch import BatchSearch
    from riptide_amd.ffautils import generate_width_trials

    dev = _lib.device()
    stream = torch.cuda.Stream()
    tsamp = 64e-6
    pmin, pmax = 0.1, 10.0
    coin = np.random.RandomState(args.seed).randint(0, 2, size=args.trials)
    bmin = np.where(coin == 0, 240, 480)
    bmax = np.where(coin == 0, 260, 520)

    with torch.cuda.stream(stream):
        t0 = time.time()
        xs = make_inputs(torch, args.trials, args.nsamp, tsamp, args.seed, dev)
        torch.cuda.synchronize()
        log(f"inputs: {args.trials} x {args.nsamp} samples in {time.time() - t0:.1f} s")

        batch = BatchSearch([args.nsamp] * args.trials, tsamp, bmin, bmax, pmin, pmax,
                            deredden=True, rmed_width=4.0, already_normalised=False,
                            group_size=args.group_size)
        outs = batch.allocate_outputs()
        batch.bind(xs, outs)
        st = batch.stats()
        log(f"batch: {st}")
        use_graph = not args.no_graph

        for _ in range(args.warmup):
            batch.run(use_graph=use_graph)
        batch.check()
        torch.cuda.synchronize()

        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            batch.run(use_graph=use_graph)
        ev1.record(stream)
        torch.cuda.synchronize()
        batched_ms = ev0.elapsed_time(ev1) / args.steps
        batch.check()
        trials_per_s = args.trials / (batched_ms * 1e-3)
        log(f"batched: {batched_ms:.1f} ms per batch -> {trials_per_s:.1f} trials/s")

        # per-phase event timing of one eager run (same kernels, same stream)
        phases = batch.run_timed()
        log(f"phases: {phases}")

        if args.batch_only:
            print(json.dumps({"workload": f"batched ffa_search, {args.trials} x {args.nsamp}",
                              "trials_per_s": round(trials_per_s, 2),
                              "ms_per_step": round(batched_ms, 3), "phases_ms": phases,
                              "stats": st}), flush=True)
            return

        # the S/N stay in HBM; what copying them to the host would cost
        d2h = d2h_timing(torch, outs)
        log(f"S/N device-to-host copy: {d2h}")

        # peak detection on the batch's S/N, left on the device (the consumer
        # of the periodograms)
        peaks_info, peaks_ctx = peak_timing(torch, batch, outs, bmin, tsamp, args.nsamp)
        log(f"find_peaks: {peaks_info}")

        # plain schedule: one series at a time, per-step launches, no graph.
        # Timed per bins class on the first plain_trials/2 series of each
        # class, then weighted by the batch's class mix, so that both rates
        # describe the same workload
        per_class = max(1, min(args.plain_trials, args.trials) // 2)
        widths = {b: generate_width_trials(b) for b in (240, 480)}
        tmp = _lib.empty((args.nsamp,))
        tmp2 = _lib.empty((args.nsamp,))
        plain_out = {b: libcpp.periodogram_device(tmp, args.nsamp, tsamp, widths[b], pmin, pmax,
                                                  b, b + (20 if b == 240 else 40))
                     for b in (240, 480)}  # warm-up + buffers
        torch.cuda.synchronize()
        ws = int(round(4.0 / tsamp))
        plain_ms = {}
        plain_shape = {}

        def plain_run(idx):
            for i in idx:
                b0, b1 = int(bmin[i]), int(bmax[i])
                _lib.call("rt_deredden", _lib.ptr(xs[i]), args.nsamp, ws, 101, _lib.ptr(tmp),
                          _lib.stream())
                _lib.call("rt_normalise", _lib.ptr(tmp), args.nsamp, _lib.ptr(tmp2),
                          _lib.stream())
                P, F, S = plain_out[b0]
                w = np.ascontiguousarray(widths[b0], dtype=np.uint64)
                _lib.call("rt_periodogram_plain", _lib.ptr(tmp2), args.nsamp, tsamp,
                          _lib.host_ptr(w), w.size, pmin, pmax, b0, b1, _lib.ptr(P),
                          _lib.ptr(F), _lib.ptr(S), _lib.stream())

        # the plain schedule gets its best FFA workgroup shape per class
        # (one untimed series per shape first)
        for cls in (0, 1):
            idx = [i for i in range(args.trials) if coin[i] == cls][:per_class]
            best = None
            for shape in range(tune.NUM_SHAPES):
                _lib.call("rt_set_ffa_shape", shape)
                plain_run(idx[:1])
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                plain_run(idx)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) / max(1, len(idx)) * 1e3
                if best is None or ms < best[0]:
                    best = (ms, shape)
            _lib.call("rt_set_ffa_shape", -1)
            plain_ms[cls], plain_shape[cls] = best
        frac480 = float((coin == 1).mean())
        plain_trial_ms = (1 - frac480) * plain_ms[0] + frac480 * plain_ms[1]
        plain_tps = 1e3 / plain_trial_ms
        nplain = 2 * per_class
        log(f"plain: {plain_ms[0]:.2f} ms per 240-260 trial, {plain_ms[1]:.2f} ms per 480-520 "
            f"trial -> {plain_tps:.1f} trials/s at the batch's class mix")

        # the drop-in one-series path (libcpp.periodogram, what ffa_search
        # calls): a kept batch of one series per call; same series, same
        # dereddening and normalisation per series, one untimed call per
        # class first (it builds the kept batch)
        dropin_ms = {}
        for cls in (0, 1):
            idx = [i for i in range(args.trials) if coin[i] == cls][:per_class]

            def dropin_run(ids):
                for i in ids:
                    b0, b1 = int(bmin[i]), int(bmax[i])
                    _lib.call("rt_deredden", _lib.ptr(xs[i]), args.nsamp, ws, 101, _lib.ptr(tmp),
                              _lib.stream())
                    _lib.call("rt_normalise", _lib.ptr(tmp), args.nsamp, _lib.ptr(tmp2),
                              _lib.stream())
                    libcpp.periodogram_device(tmp2, args.nsamp, tsamp, widths[b0], pmin, pmax,
                                              b0, b1)

            dropin_run(idx[:1])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dropin_run(idx)
            torch.cuda.synchronize()
            dropin_ms[cls] = (time.perf_counter() - t0) / max(1, len(idx)) * 1e3
        dropin_tps = 1e3 / ((1 - frac480) * dropin_ms[0] + frac480 * dropin_ms[1])
        log(f"drop-in: {dropin_ms[0]:.2f} ms per 240-260 trial, {dropin_ms[1]:.2f} ms per "
            f"480-520 trial -> {dropin_tps:.1f} trials/s at the batch's class mix")

    # heterogeneous batch: 512 series, lengths log-uniform 2^21..2^23 rounded
    # to a multiple of 1024, same noise model and bins coin
    c3 = None
    if not args.no_c3:
        rs3 = np.random.RandomState(args.seed + 3)
        lens = (np.exp(rs3.uniform(np.log(2 ** 21), np.log(2 ** 23), size=args.trials))
                .astype(np.int64) // 1024 * 1024)
        with torch.cuda.stream(stream):
            t0 = time.time()
            x3 = make_inputs(torch, args.trials, lens, tsamp, args.seed + 3, dev)
            torch.cuda.synchronize()
            log(f"heterogeneous inputs: {args.trials} series, {int(lens.sum())} samples in "
                f"{time.time() - t0:.1f} s")
            b3 = BatchSearch(lens, tsamp, bmin, bmax, pmin, pmax, deredden=True, rmed_width=4.0,
                             already_normalised=False, group_size=args.group_size)
            o3 = b3.allocate_outputs()
            b3.bind(x3, o3)
            for _ in range(args.warmup):
                b3.run(use_graph=use_graph)
            torch.cuda.synchronize()
            ev0.record(stream)
            for _ in range(args.steps):
                b3.run(use_graph=use_graph)
            ev1.record(stream)
            torch.cuda.synchronize()
            b3.check()
            ms3 = ev0.elapsed_time(ev1) / args.steps
        c3 = {"workload": (f"{args.trials} series, lengths log-uniform 2^21-2^23 (multiple of "
                           f"1024, seeded), same noise, bins coin and search as the headline"),
              "ms_per_batch": round(ms3, 3),
              "trials_per_s": round(args.trials / (ms3 * 1e-3), 2),
              "equiv_2p22_trials_per_s": round(float(lens.sum()) / 2 ** 22 / (ms3 * 1e-3), 2),
              "total_samples": int(lens.sum())}
        log(f"heterogeneous batch: {c3}")
        b3.close()
        del x3, o3, b3

    # config 2: one 2^23-sample series through the drop-in path
    with torch.cuda.stream(stream):
        c2 = c2_timing(torch, _lib)
    c2_info = {"workload": ("one series of 2^23 samples at 64 us, white + red noise, pulses at "
                            "0.2371 s and 4.9133 s; dereddening, normalisation, periodogram P "
                            "0.1-10 s, bins 480-520 (2,652,094 periods x 12 widths)"),
               "gpu_ms": round(c2[5], 3)}
    log(f"config 2: {c2_info}")
    # config 1 series for the float64 error statement (GPU S/N now, the
    # recomputation in the CPU leg)
    from riptide_amd import synthetic
    from riptide_amd.search import ffa_search
    from riptide_amd.time_series import TimeSeries
    kw1 = dict(period_min=0.5, period_max=2.0, bins_min=240, bins_max=260)
    x1 = synthetic.series(2 ** 20, 256e-6, seed=1, red_index=None,
                          pulses=[(1.3711, 0.02, 20.0)])
    _, pg1 = ffa_search(TimeSeries(x1, 256e-6), rmed_width=4.0, **kw1)
    c1 = (x1, 256e-6, kw1, pg1.snrs)

    # differentiable search (config 4 shape): forward vs forward + adjoint
    grad = fwd_bwd_timing(torch, _lib)
    log(f"fwd+bwd: {grad}")
    c5 = c5_timing(torch, _lib, args.seed, adam_steps=args.c5_steps) if args.c5_steps > 0 else None

    # roofline of the dominant kernel: persistent FFA + S/N kernel
    snr_bytes = sum(batch.lengths[i] * batch.widths[i].size * 4 for i in range(args.trials))
    fold_bytes = st["fold_floats"] * 4
    # the unit of the roofline figures is one group of series: its FFA
    # launches, one per kernel variant (row-length class) present in it
    groups = max(1, (args.trials + args.group_size - 1) // args.group_size)
    launches = max(1, phases["ffa_launches"])
    bytes_per_group = (fold_bytes + snr_bytes) / groups
    ms_per_group = phases["ffa_ms"] / groups
    achieved = bytes_per_group / (ms_per_group * 1e-3) / 1e9
    lds_bytes = 8 * st["ffa_adds"]
    lds_achieved = lds_bytes / (phases["ffa_ms"] * 1e-3) / 1e9

    # HBM bytes per launch of the same kernel and workload, measured with two
    # rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by a copy-kernel
    # calibration; committed under profiles/ (the counters cannot be read
    # from inside this process)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ffa_batch_kernel__batch512x2p22_traffic.json")
    if (os.path.exists(tpath) and args.trials == 512 and args.nsamp == 2 ** 22
            and args.group_size == 64):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = int(tj["hbm_bytes_per_group"]) if "hbm_bytes_per_group" in tj else None

    # CPU-baseline leg: the oracle timed on the host cores, and the checks of
    # the GPU results made with the same oracle
    cpu = checks = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline(args.nsamp, tsamp)
        log(f"cpu baseline: {cpu} ({time.time() - t0:.1f} s)")
        t0 = time.time()
        checks = oracle_checks(batch, xs, outs, coin, bmin, bmax, tsamp, pmin, pmax, c1, c2,
                               peaks_ctx)
        log(f"oracle checks: {checks} ({time.time() - t0:.1f} s)")
        c2_info.update(checks.pop("c2"))
        peaks_info["host_path_ms_per_trial"] = checks.pop("find_peaks_host_ms_per_trial")
        peaks_info["same_peaks_as_host_path"] = checks.pop("find_peaks_same_as_host_path")

    line = {
        "metric": "DM trials/sec (2^22-sample ffa_search), batched vs plain, 1 GPU; fwd+bwd ms",
        "value": round(trials_per_s, 2),
        "unit": "DM trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (white + red noise, 1 in 16 with a pulse), generated on the GPU",
        "config": {
            "workload": (f"batched ffa_search: {args.trials} DM trials x {args.nsamp} samples, "
                         f"tsamp 64 us, P 0.1-10 s, bins (240,260)/(480,520) by seeded coin, "
                         f"dereddening 4 s + normalisation, S/N left in HBM"),
            "trials": args.trials,
            "nsamp": args.nsamp,
            "bins_classes": {"240-260": int((coin == 0).sum()), "480-520": int((coin == 1).sum())},
            "hipgraph": use_graph,
            "group_size": args.group_size,
        },
        "plain": {"value": round(plain_tps, 2), "unit": "DM trials/s", "trials": nplain,
                  "ms_per_trial_by_class": {"240-260": round(plain_ms[0], 3),
                                            "480-520": round(plain_ms[1], 3)},
                  "workgroup_shape_by_class": {"240-260": plain_shape[0],
                                               "480-520": plain_shape[1]},
                  "schedule": ("one series per launch sequence: per-step launches, no queue, no "
                               "graph; timed per bins class under every FFA workgroup shape, "
                               "the fastest kept, and weighted by the batch's class mix")},
        "batched_over_plain": round(trials_per_s / plain_tps, 3),
        "dropin": {"value": round(dropin_tps, 2), "unit": "DM trials/s", "trials": nplain,
                   "ms_per_trial_by_class": {"240-260": round(dropin_ms[0], 3),
                                             "480-520": round(dropin_ms[1], 3)},
                   "schedule": ("the drop-in one-series path (libcpp.periodogram, as ffa_search "
                                "calls it): a kept batch of one series per call, eager launches; "
                                "same series, dereddening and normalisation as the plain line")},
        "batched_over_dropin": round(trials_per_s / dropin_tps, 3),
        "phases_ms": {k: round(v, 3) if isinstance(v, float) else v for k, v in phases.items()},
        "parity": checks["parity"] if checks else None,
        "snr_vs_fp64": checks["snr_vs_fp64"] if checks else None,
        "snr_d2h": d2h,
        "c2_single_series": c2_info,
        "roofline": {
            "bound": "hbm",
            "binding_resource": ("LDS bandwidth and latency (merge levels, S/N windows), not "
                                 "HBM: see roofline.lds; HBM is the contract figure"),
            "kernel": ("ffa_batch_kernel (fold + FFA merge levels + boxcar S/N); figures per "
                       "group of series: its launches, one per kernel variant"),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": ("profiles/ffa_batch_kernel__batch512x2p22_traffic.json"
                               if traffic is not None else None),
            "algorithmic_bytes_per_group": int(bytes_per_group),
            "ms_per_group": round(ms_per_group, 3),
            "launches_per_group": round(launches / groups, 3),
            "lds": {"achieved": round(lds_achieved, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
                    "frac": round(lds_achieved / LDS_PEAK_GBS, 4),
                    "peak_measured": round(LDS_MEASURED_GBS, 1),
                    "frac_measured": round(lds_achieved / LDS_MEASURED_GBS, 4),
                    "bytes": "8 B per FFA float add",
                    "peaks": ("nominal 128 B/clk/CU; measured ds_read_b32 rate of gfx950 "
                              "169 B/clk/CU (profiles/lds_read_rates__gfx950.md); 256 CU at "
                              "2.4 GHz")},
        },
        "heterogeneous_batch": c3,
        "find_peaks": peaks_info,
        "fwd_bwd": grad,
        "fir_whitening_batch": c5,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
