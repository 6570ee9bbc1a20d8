/*
 * riptide_amd C-ABI: the drop-in boundary of the riptide FFA periodogram path
 * on AMD Instinct MI355X (gfx950).
 *
 * Every entry point replaces one function of the reference's pybind11 module
 * `riptide.libcpp` (src/riptide/cpp/python_bindings.cpp) or one numpy step of
 * its Python layer, cited next to each declaration.  Conventions:
 *
 *   - Every function returns an int status: RT_OK, RT_EINVAL (argument error,
 *     raised as ValueError with the reference's message, see rt_last_error)
 *     or RT_EDEVICE (HIP error, raised as RuntimeError).
 *   - Array arguments named d_* are DEVICE pointers (HBM).  Arguments named
 *     h_* are host pointers.  Sizes are int64_t element counts.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Unless stated
 *     otherwise a call returns once its work is complete on `stream`.
 *   - Outputs are caller-allocated; the library keeps a per-device scratch
 *     pool for intermediate buffers, so concurrent calls on one device are
 *     serialised internally.
 */
#ifndef RIPTIDE_AMD_H
#define RIPTIDE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_EINVAL 1
#define RT_EDEVICE 2

/* Message of the last failed call on this thread. */
const char* rt_last_error(void);

/* Library ABI version. */
int rt_version(void);

/* Workgroup shape of the FFA kernels for subsequent calls: 0 = one
 * 1024-thread workgroup per CU (~159 KiB LDS), 1 = two 512-thread workgroups
 * per CU (~79 KiB each), -1 = the tuned default.  Results are identical;
 * only speed differs (riptide_amd/tune.py picks the faster). */
int rt_set_ffa_shape(int shape);

/* Releases the device memory the library keeps between calls on the current
 * device: the scratch pool of the pass-by-pass schedules and the kept batches
 * of the last one-series searches (rt_periodogram).  No counterpart in the
 * reference, which allocates inside every call; for callers that need that
 * memory back between searches.  Waits for the device. */
int rt_release_scratch(void);

/* ---- Host-side search plan (no device work) ------------------------------ */

/* periodogram_length (periodogram.hpp:63-109; python_bindings.cpp:196):
 * number of trial periods of a search. */
int rt_periodogram_length(int64_t size, double tsamp, double period_min, double period_max,
                          int64_t bins_min, int64_t bins_max, int64_t* length);

/* Trial periods and fold bins of a search, host arrays of `length` elements
 * (periodogram.hpp:190-194). */
int rt_periodogram_grid(int64_t size, double tsamp, double period_min, double period_max,
                        int64_t bins_min, int64_t bins_max, double* h_periods,
                        uint32_t* h_foldbins);

/* Index bounds and edge weights of downsampled sample k (downsample.hpp:58-73),
 * computed on the host by the same exact integer emulation of the float64
 * arithmetic that the downsampling kernels run. */
int rt_downsample_index(int64_t k, double f, int64_t size, int64_t* imin, int64_t* imax,
                        float* wmin, float* wmax);

/* ---- Kernels of riptide.libcpp (python_bindings.cpp:225-279) -------------- */

/* libcpp.ffa2 (python_bindings.cpp:84-96 -> transforms.hpp:53-61):
 * FFA transform of a rows x cols float32 block; bit-identical output. */
int rt_ffa2(const float* d_in, int64_t rows, int64_t cols, float* d_out, void* stream);

/* libcpp.snr2 / snr1 (python_bindings.cpp:121-160 -> snr.hpp:37-65):
 * boxcar S/N of rows x cols profiles for nw trial widths; d_out is rows x nw. */
int rt_snr2(const float* d_in, int64_t rows, int64_t cols, const uint64_t* h_widths, int64_t nw,
            float stdnoise, float* d_out, void* stream);

/* libcpp.downsample (python_bindings.cpp:163-177 -> downsample.hpp:44-82):
 * d_out holds floor(size / f) samples. */
int rt_downsample(const float* d_in, int64_t size, double f, float* d_out, void* stream);

/* libcpp.periodogram (python_bindings.cpp:180-209 -> periodogram.hpp:117-201):
 * d_periods / d_foldbins / d_snrs hold length, length and length x nw values.
 * Runs as a batch of one series (rt_batch_*, with the batch tiling table in
 * effect); the batches of the last four searches are kept for later calls
 * with one of those searches.  Waits for the stream. */
int rt_periodogram(const float* d_in, int64_t size, double tsamp, const uint64_t* h_widths,
                   int64_t nw, double period_min, double period_max, int64_t bins_min,
                   int64_t bins_max, double* d_periods, uint32_t* d_foldbins, float* d_snrs,
                   void* stream);

/* The same search run as one launch sequence per downsampling step (the
 * "plain" schedule); identical results, kept for timing comparisons. */
int rt_periodogram_plain(const float* d_in, int64_t size, double tsamp, const uint64_t* h_widths,
                         int64_t nw, double period_min, double period_max, int64_t bins_min,
                         int64_t bins_max, double* d_periods, uint32_t* d_foldbins,
                         float* d_snrs, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rt_periodogram with every intermediate buffer carved from
 * a static device array of 256 MiB per device instead of the library's
 * scratch pool; no hipMalloc / hipFree.  Identical results.  Searches whose
 * intermediates exceed the array (roughly: packed downsampled series, the
 * largest transform's two intermediate states and 24 MiB of task list) are
 * rejected with RT_EINVAL, as are rows wider than the staging of the
 * workgroup shape in effect (it runs pass by pass in that shape, where
 * rt_periodogram tiles such rows with a deeper shape).  Calls are serialised by the library's lock and
 * wait for their work; the array keeps the last call's contents, and calls
 * that overlap on one GPU from several processes each use their own copy. */
int rt_periodogram_noalloc(const float* d_in, int64_t size, double tsamp,
                           const uint64_t* h_widths, int64_t nw, double period_min,
                           double period_max, int64_t bins_min, int64_t bins_max,
                           double* d_periods, uint32_t* d_foldbins, float* d_snrs, void* stream);

/* libcpp.running_median (python_bindings.cpp:212-222 -> running_median.hpp:100-132). */
int rt_running_median(const float* d_in, int64_t size, int64_t width, float* d_out,
                      void* stream);

/* libcpp.rollback (python_bindings.cpp:44-52 -> kernels.hpp:32-38). */
int rt_rollback(const float* d_in, int64_t size, int64_t shift, float* d_out, void* stream);

/* libcpp.fused_rollback_add (python_bindings.cpp:55-70 -> kernels.hpp:19-25). */
int rt_fused_rollback_add(const float* d_x, const float* d_y, int64_t size, int64_t shift,
                          float* d_out, void* stream);

/* libcpp.circular_prefix_sum (python_bindings.cpp:73-81 -> kernels.hpp:73-101).
 * d_out holds nsum elements (the reference allocates only `size`). */
int rt_circular_prefix_sum(const float* d_in, int64_t size, int64_t nsum, float* d_out,
                           void* stream);

/* libcpp.benchmark_ffa2 (python_bindings.cpp:99-118): seconds per FFA of a
 * zero rows x cols block, timed on the device. */
int rt_benchmark_ffa2(int64_t rows, int64_t cols, int64_t loops, double* seconds);

/* Device milliseconds of one FFA transform of a rows x cols block (with the
 * fused S/N of every row when with_snr) using workgroup shape `shape`,
 * averaged over `reps` launches; used by the tile autotuner. */
int rt_time_ffa(int64_t rows, int64_t cols, int shape, int with_snr, int64_t reps, double* ms);

/* ---- Python-side numerics of the path (numpy in the reference) ----------- */

/* scrunch (running_medians.py:40-46): float32 means of blocks of `factor`
 * samples; d_out holds size / factor values. */
int rt_scrunch_mean(const float* d_in, int64_t size, int64_t factor, float* d_out, void* stream);

/* TimeSeries.deredden (time_series.py:95-124 with fast_running_median,
 * running_medians.py:49-85): d_out = data - approximate running median. */
int rt_deredden(const float* d_in, int64_t size, int64_t width_samples, int64_t min_points,
                float* d_out, void* stream);

/* fast_running_median (running_medians.py:49-85) alone, as float32. */
int rt_fast_running_median(const float* d_in, int64_t size, int64_t width_samples,
                           int64_t min_points, float* d_out, void* stream);

/* TimeSeries.normalise (time_series.py:67-93): zero mean, unit variance. */
int rt_normalise(const float* d_in, int64_t size, float* d_out, void* stream);

/* ---- Adjoints (no counterpart in the reference, which has no gradients) -- */

/* rt_normalise that also returns h_stats[4] = mean (hi, lo), std (hi, lo). */
int rt_normalise_ex(const float* d_in, int64_t size, float* d_out, float* h_stats, void* stream);

/* Transpose of normalisation: d_gx = (g - mean(g) - y*mean(g*y)) * inv_std. */
int rt_normalise_backward(const float* d_g, const float* d_y, int64_t size, float inv_std,
                          float* d_gx, void* stream);

/* Transpose of the FFA transform (rt_ffa2): d_gin = A^T d_gout. */
int rt_ffa2_backward(const float* d_gout, int64_t rows, int64_t cols, float* d_gin, void* stream);

/* Gradient of sum(gsnr * snr2(x)) with respect to x (rows x cols), through
 * the first argmax window of every width. */
int rt_snr2_backward(const float* d_in, int64_t rows, int64_t cols, const uint64_t* h_widths,
                     int64_t nw, float stdnoise, const float* d_gsnr, float* d_gin, void* stream);

/* Gradient of sum(gsnr * periodogram(x)) with respect to the (normalised)
 * series x, recomputing the forward of every (step, bins) transform whose
 * rows receive gradient: S/N backward -> FFA transpose -> fold transpose ->
 * downsampling transpose.  d_gsnr is length x nw, d_gin has size floats. */
int rt_periodogram_backward(const float* d_in, int64_t size, double tsamp,
                            const uint64_t* h_widths, int64_t nw, double period_min,
                            double period_max, int64_t bins_min, int64_t bins_max,
                            const float* d_gsnr, float* d_gin, void* stream);

/* ---- Batched differentiable periodogram (no counterpart in the reference,
 * which has no gradients): nseries series of `size` samples that share one
 * search plan, stored row-major (series i at d + i*size).  The forward
 * periodogram of the normalised rows runs as a batch (rt_batch_*); these
 * entry points normalise the rows and take the gradient back to the
 * un-normalised input in a fixed number of launches for the whole batch.
 * Results are bit-identical to rt_normalise_ex + rt_periodogram_backward +
 * rt_normalise_backward on each series. ---------------------------------- */

/* Plans the batch (widths 1 to 64 of them, each > 0 and < bins_min). */
int rt_grad_batch_create(int64_t nseries, int64_t size, double tsamp, const uint64_t* h_widths,
                         int64_t nw, double period_min, double period_max, int64_t bins_min,
                         int64_t bins_max, void** handle);

/* Trial periods per series. */
int rt_grad_batch_length(void* handle, int64_t* length);

/* d_y = every row of d_x normalised (time_series.py:67-93); h_inv_std (may
 * be NULL) receives 1/std of every row, in float32 as the one-series
 * adjoint takes it.  Waits for the work. */
int rt_grad_batch_normalise(void* handle, const float* d_x, float* d_y, float* h_inv_std,
                            void* stream);

/* d_gx (nseries x size) = gradient of sum(d_gsnr * periodogram(d_y)) with
 * respect to the un-normalised rows, given the normalised rows d_y and their
 * h_inv_std; d_gsnr is nseries x length x nw.  Recomputes the forward of
 * every (series, transform) pair that receives gradient.  Waits. */
int rt_grad_batch_backward(void* handle, const float* d_y, const float* h_inv_std,
                           const float* d_gsnr, float* d_gx, void* stream);

int rt_grad_batch_destroy(void* handle);

/* ---- Batched DM-trial search (no counterpart in the reference: it replaces
 * the reference's one-process-per-DM-trial pool, pipeline/worker_pool.py:35-45,
 * each worker calling ffa_search, search.py:10-93) -------------------------- */

/* Plans a batch of nseries searches.  Series i has sizes[i] samples, sampling
 * time tsamps[i], bins range [bins_min[i], bins_max[i]] and widths_count[i]
 * trial widths (consecutive in widths_flat).  deredden / normalise select the
 * preparation of ffa_search (search.py:82-86).  Series are processed in
 * groups of group_size.  Allocates every device buffer the batch needs. */
int rt_batch_create(int64_t nseries, const int64_t* sizes, const double* tsamps,
                    const uint64_t* widths_flat, const int64_t* widths_count, double period_min,
                    double period_max, const int64_t* bins_min, const int64_t* bins_max,
                    int deredden, double rmed_width, int64_t rmed_minpts, int normalise,
                    int64_t group_size, void** handle);

/* Tiling of the batch's transforms by (rows, bins) class, for batches created
 * afterwards: n entries of 7 int32 each, {rows_lo, rows_hi, bins_lo, bins_hi,
 * shape, cap, lmax}; a transform of r rows and b bins takes the first entry
 * with rows_lo <= r < rows_hi and bins_lo <= b < bins_hi: FFA workgroup
 * shape, rows per tile (0: as many as fit) and merge levels per cone pass
 * (0: the planner's limit).  Transforms no entry holds take the current
 * shape (rt_set_ffa_shape) with the largest tile.  n = 0 clears the table. */
int rt_batch_set_tiling(const int32_t* entries, int n);

/* Number of trial periods of series i, and its period / fold-bins grid. */
int rt_batch_length(void* handle, int64_t i, int64_t* length);
int rt_batch_grid(void* handle, int64_t i, double* h_periods, uint32_t* h_foldbins);

/* Binds device input series and S/N outputs (length x nw floats each); host
 * arrays of nseries device pointers.  Synchronous. */
int rt_batch_bind(void* handle, const float* const* h_inputs, float* const* h_snrs);

/* Enqueues the whole batch on `stream` without host synchronisation.  With
 * use_graph the launch sequence is captured once into a hipGraph and
 * replayed on later calls (until the next rt_batch_bind). */
int rt_batch_run(void* handle, int use_graph, void* stream);

/* One eager run with HIP events around each phase of each group, on
 * `stream`; waits.  ms[0..3] = preparation ms, downsampling ms, FFA + S/N ms,
 * number of FFA launches (one per group and FFA kernel variant). */
int rt_batch_run_timed(void* handle, void* stream, double* ms);

/* Waits for `stream` and reports a device-side tiling error, if any. */
int rt_batch_check(void* handle, void* stream);

/* stats[10]: transforms, groups, workgroup slots, scratch bytes, packed
 * downsampling bytes, folded input floats, trial periods, FFA float adds,
 * tile records (the cones and row tables of transforms several series share,
 * built once per run) and their bytes. */
int rt_batch_stats(void* handle, int64_t* stats);

int rt_batch_destroy(void* handle);

/* ---- FIR whitening front end of a differentiable batched search (no
 * counterpart in the reference, which dereddens with a running median only,
 * time_series.py:95-124; the filters are optimised for S/N by gradient
 * descent through rt_periodogram_backward) ----------------------------------- */

/* nseries series of size samples each (row-major, d_x), each with its own odd
 * number ntaps of taps (row-major, d_taps), centre c = (ntaps - 1)/2, zero
 * outside the series: y[s][n] = sum_k taps[s][k] * x[s][n + k - c].  With
 * reverse the taps are applied in reverse order, which gives the gradient
 * with respect to x of sum(g * y) from g. */
int rt_fir_forward(const float* d_x, int64_t nseries, int64_t size, const float* d_taps,
                   int64_t ntaps, int reverse, float* d_y, void* stream);

/* Floats of device workspace rt_fir_grad_taps needs. */
int rt_fir_workspace_floats(int64_t nseries, int64_t size, int64_t ntaps, int64_t* floats);

/* Gradient of sum(g * y) with respect to the taps:
 * gtaps[s][k] = sum_n g[s][n] * x[s][n + k - c] (nseries x ntaps floats). */
int rt_fir_grad_taps(const float* d_x, const float* d_g, int64_t nseries, int64_t size,
                     int64_t ntaps, float* d_work, float* d_gtaps, void* stream);

/* One Adam step (Kingma & Ba 2015) on count parameters, step counted from 1;
 * d_m and d_v hold the moment estimates (zero before the first step). */
int rt_adam_update(float* d_params, const float* d_grad, float* d_m, float* d_v, int64_t count,
                   double lr, double beta1, double beta2, double eps, int64_t step, void* stream);

/* ---- Batched input staging (reading/presto.py, reading/sigproc.py and
 * time_series.py:301-377 read one file at a time into host float32 arrays;
 * here files are read into one pinned buffer, copied to the device in one
 * transfer and 8-bit samples widened there) -------------------------------- */

/* n 8-bit samples (unsigned, or signed when is_signed) at d_raw to float32. */
int rt_convert_samples(const void* d_raw, int64_t n, int nbits, int is_signed, float* d_out,
                       void* stream);

/* ---- Candidate folding (folding.py:21-85; the reference folds on the host
 * through libcpp.downsample) ------------------------------------------------ */

/* The rows x bins block of a downsampled series (d_x, modified in place) is
 * scaled by the float32 factor `scale`; with sum, d_out[b] receives the sum of
 * the scaled rows (added row after row). */
int rt_fold_rows(float* d_x, int64_t rows, int64_t bins, float scale, int sum, float* d_out,
                 void* stream);

/* downsample_vertical (folding.py:6-18): every column of a rows x bins array
 * downsampled by f (1 < f < rows) into floor(rows / f) rows. */
int rt_downsample_vertical(const float* d_x, int64_t rows, int64_t bins, double f, float* d_out,
                           void* stream);

/* ---- Peak detection on the device (peak_detection.py:38-240: find_peaks
 * runs on host copies of the S/N array in the reference; here only the
 * segment statistics and the candidate indices leave the device) ---------- */

/* d_snrs: length x nw S/N (row-major).  For each of nseg consecutive segments
 * of seglen rows and each width: the values of the given 0-based ranks
 * (ascending order) within the segment, d_out[(seg * nw + w) * nranks + t].
 * Exact order statistics (radix select); at most 8 ranks. */
int rt_segment_select(const float* d_snrs, int64_t nw, int64_t nseg, int64_t seglen,
                      const int32_t* h_ranks, int64_t nranks, float* d_out, void* stream);

/* Peak candidates (peak_detection.py:154-158): rows ip where snr[ip][w] > smin
 * (smin_lo: the largest float32 <= smin) and snr[ip][w] > poly_w(log f[ip]),
 * d_coef holding ncoef float32 coefficients per width (highest degree first),
 * d_logf float32 log-frequencies.  Sure passes are appended to d_sel[w * cap
 * + ...], values within the float32 evaluation's error bound of the
 * threshold to d_unc[w * cap + ...] (the caller decides those in float64);
 * d_counts[2w] and d_counts[2w + 1] count them (must be zero on entry; may
 * exceed cap, in which case the lists are incomplete). */
int rt_peak_candidates(const float* d_snrs, int64_t length, int64_t nw, const float* d_logf,
                       const float* d_coef, int64_t ncoef, float smin_lo, int32_t* d_counts,
                       int32_t* d_sel, int32_t* d_unc, int64_t cap, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RIPTIDE_AMD_H */
