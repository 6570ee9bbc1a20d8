/*
 * CPU oracle for the riptide FFA periodogram path.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test-suite, the
 * smoke check and the CPU-baseline leg of bench.py may load it, and only as
 * the checker (or as the timed CPU baseline).  The product path in
 * riptide_amd/ never links, loads or calls anything in this directory.
 *
 * It restates, in plain C and in the reference's operation order, the native
 * routines of riptide v0.2.5 (src/riptide/cpp/*.hpp).  Every function cites
 * the reference lines it follows.  Compiled with the reference's own
 * "safe fast-math" flags (setup.py:13-20,34) so that compiler-dependent
 * choices (contraction, vectorised reductions) match a reference build made
 * with the same compiler and -march.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   ffa transform  : 8x8 known-answer test, integer brute force, linearity,
 *                    bit-exact vs the compiled reference (oracle/_ref).
 *   snr            : top-hat closed form, rotation invariance, float64
 *                    brute-force boxcar, bit-exact vs oracle/_ref.
 *   downsample     : integer-factor block sums, float64 integration of the
 *                    step function, bit-exact vs oracle/_ref.
 *   running median : naive edge-padded np.median, bit-exact vs oracle/_ref.
 *   periodogram    : plan length / period grid closed forms, bit-exact vs
 *                    oracle/_ref.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "riptide_oracle.h"

static __thread char g_err[256];

const char* or_last_error(void) { return g_err; }

static int fail(const char* msg)
{
    strncpy(g_err, msg, sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
    return OR_EINVAL;
}

/* ------------------------------------------------------------------------ */
/* Elementwise helpers: kernels.hpp:9-46                                     */
/* ------------------------------------------------------------------------ */

/* z = x + y (kernels.hpp:9-13) */
static void vadd(const float* restrict x, const float* restrict y, size_t n, float* restrict z)
{
    for (size_t i = 0; i < n; ++i)
        z[i] = x[i] + y[i];
}

/* z = x + roll(y, -shift) (kernels.hpp:19-25): the first size-p outputs add
 * y[p..size), the last p outputs add y[0..p). */
void or_fused_rollback_add(const float* x, const float* y, size_t size, size_t shift, float* z)
{
    const size_t p = shift % size;
    vadd(x, y + p, size - p, z);
    vadd(x + (size - p), y, p, z + (size - p));
}

/* out = roll(x, -shift) (kernels.hpp:32-38) */
void or_rollback(const float* x, size_t size, size_t shift, float* out)
{
    const size_t p = shift % size;
    memcpy(out, x + p, (size - p) * sizeof(float));
    memcpy(out + (size - p), x, p * sizeof(float));
}

/* max_i (x[i] - y[i]) (kernels.hpp:50-60) */
static float diff_max(const float* restrict x, const float* restrict y, size_t n)
{
    float best = x[0] - y[0];
    for (size_t i = 1; i < n; ++i) {
        const float d = x[i] - y[i];
        if (d > best)
            best = d;
    }
    return best;
}

/* Circular prefix sum with a float64 accumulator over the first wrap, then
 * float32 "out[j] + i*sum" for the following wraps (kernels.hpp:73-101).
 * out must hold nsum elements. */
void or_circular_prefix_sum(const float* x, size_t size, size_t nsum, float* out)
{
    double acc = 0.0;
    const size_t first = size < nsum ? size : nsum;
    for (size_t j = 0; j < first; ++j) {
        acc += x[j];
        out[j] = (float)acc;
    }
    if (nsum <= size)
        return;
    const float total = (float)acc;
    const size_t wraps = nsum / size;
    const size_t rest = nsum % size;
    for (size_t i = 1; i < wraps; ++i) {
        const float a = (float)i * total;
        for (size_t j = 0; j < size; ++j)
            out[i * size + j] = out[j] + a;
    }
    const float a = (float)wraps * total;
    for (size_t j = 0; j < rest; ++j)
        out[wraps * size + j] = out[j] + a;
}

/* ------------------------------------------------------------------------ */
/* FFA transform: transforms.hpp:13-61, block.hpp:30-43                      */
/* ------------------------------------------------------------------------ */

/* Merge the transforms of the head (first m>>1 rows) and of the tail into
 * out (transforms.hpp:13-27).  Row s adds head row h(s) and tail row t(s)
 * rolled left by s - t(s); h and t are rounded in float32. */
static void ffa_merge(const float* head, size_t mh, const float* tail, size_t mt,
                      float* out, size_t m, size_t p)
{
    const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
    const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
    for (size_t s = 0; s < m; ++s) {
        const size_t h = (size_t)(kh * (float)s + 0.5f);
        const size_t t = (size_t)(kt * (float)s + 0.5f);
        or_fused_rollback_add(head + h * p, tail + t * p, p, s - t, out + s * p);
    }
}

/* Recursive FFA of an m x p block (transforms.hpp:30-50).  The head and tail
 * halves are transformed into temp (with out as their scratch), then merged
 * into out. */
static void ffa_rec(const float* in, float* temp, float* out, size_t m, size_t p)
{
    if (m == 2) {
        vadd(in, in + p, p, out);
        or_fused_rollback_add(in, in + p, p, 1, out + p);
        return;
    }
    if (m == 1) {
        memcpy(out, in, p * sizeof(float));
        return;
    }
    const size_t mh = m >> 1;
    const size_t mt = m - mh;
    ffa_rec(in, out, temp, mh, p);
    ffa_rec(in + mh * p, out + mh * p, temp + mh * p, mt, p);
    ffa_merge(temp, mh, temp + mh * p, mt, out, m, p);
}

int or_ffa2(const float* in, size_t rows, size_t cols, float* temp, float* out)
{
    /* The reference recurses forever on rows == 0 and divides by zero on
     * cols == 0 (block.hpp:30-31, kernels.hpp:21); reject both. */
    if (rows == 0 || cols == 0)
        return fail("ffa2 input must have at least one row and one column");
    ffa_rec(in, temp, out, rows, cols);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Boxcar S/N: snr.hpp:14-65                                                  */
/* ------------------------------------------------------------------------ */

int or_check_stdnoise(float stdnoise)
{
    if (!(stdnoise > 0))
        return fail("stdnoise must be > 0");
    return OR_OK;
}

int or_check_trial_widths(const uint64_t* widths, size_t nw, size_t bins)
{
    for (size_t i = 0; i < nw; ++i)
        if (!(widths[i] > 0 && widths[i] < bins))
            return fail("trial widths must be all > 0 and < columns");
    return OR_OK;
}

/* S/N of one profile of `size` bins for every trial width (snr.hpp:37-55).
 * scratch must hold size + max(widths) floats. */
static void snr_profile(const float* x, size_t size, const uint64_t* widths, size_t nw,
                        float stdnoise, float* scratch, float* out)
{
    size_t wmax = widths[0];
    for (size_t i = 1; i < nw; ++i)
        if (widths[i] > wmax)
            wmax = widths[i];
    or_circular_prefix_sum(x, size, size + wmax, scratch);
    const float total = scratch[size - 1];
    for (size_t iw = 0; iw < nw; ++iw) {
        const size_t w = widths[iw];
        const float h = sqrtf((size - w) / (float)(size * w));
        const float b = w / (float)(size - w) * h;
        const float dmax = diff_max(scratch + w, scratch, size);
        out[iw] = ((h + b) * dmax - b * total) / stdnoise;
    }
}

int or_snr2(const float* x, size_t rows, size_t cols, const uint64_t* widths, size_t nw,
            float stdnoise, float* out)
{
    int st = or_check_stdnoise(stdnoise);
    if (st) return st;
    st = or_check_trial_widths(widths, nw, cols);
    if (st) return st;
    if (nw == 0 || rows == 0)
        return OR_OK;
    size_t wmax = 0;
    for (size_t i = 0; i < nw; ++i)
        if (widths[i] > wmax) wmax = widths[i];
    float* scratch = (float*)malloc((cols + wmax) * sizeof(float));
    for (size_t i = 0; i < rows; ++i)
        snr_profile(x + i * cols, cols, widths, nw, stdnoise, scratch, out + i * nw);
    free(scratch);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Fractional downsampling: downsample.hpp:11-82                             */
/* ------------------------------------------------------------------------ */

size_t or_downsampled_size(size_t n, double f)
{
    return (size_t)floor(n / f);
}

double or_downsampled_variance(size_t n, double f)
{
    const double k = floor(f);
    const double r = f - k;
    const double x = or_downsampled_size(n, f) * r;
    if (x > 1)
        return f - 1.0 / 3.0;
    return pow(k - 1, 2) + 2.0 / 3.0 * pow(x, 2) - x + 1.0;
}

int or_check_downsampling_factor(size_t n, double f)
{
    if (!((f > 1.0) & (f <= n)))
        return fail("Downsampling factor must verify: 1 < f <= size");
    return OR_OK;
}

/* out[k] = wmin*in[imin] + sum(in[imin+1 .. imax-1]) + wmax*in[imax] with
 * imin = floor(k f), imax = min(floor(k f + f), N - 1) (downsample.hpp:44-82). */
int or_downsample(const float* in, size_t size, double f, float* out)
{
    int st = or_check_downsampling_factor(size, f);
    if (st) return st;
    const size_t n = or_downsampled_size(size, f);
    for (size_t k = 0; k < n; ++k) {
        const double start = k * f;
        const double end = start + f;
        const size_t imin = (size_t)floor(start);
        const size_t imax = (size_t)fmin(floor(end), size - 1.0);
        const float wmin = (imin + 1) - start;
        const float wmax = end - imax;
        float acc = wmin * in[imin];
        for (size_t i = imin + 1; i < imax; ++i)
            acc += in[i];
        acc += wmax * in[imax];
        out[k] = acc;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Periodogram: periodogram.hpp:25-201                                       */
/* ------------------------------------------------------------------------ */

int or_periodogram_check(size_t size, double tsamp, double pmin, double pmax,
                         size_t bmin, size_t bmax)
{
    (void)size;
    if (!(tsamp > 0)) return fail("tsamp must be > 0");
    if (!(pmin > 0)) return fail("period_min must be > 0");
    if (!(pmax > pmin)) return fail("period_max must be > period_min");
    if (!(bmin > 1)) return fail("bins_min must be > 1");
    if (!(bmax >= bmin)) return fail("bins_max must be >= bins_min");
    if (!(pmin >= tsamp * bmin)) return fail("Must have: period_min >= tsamp * bins_min ");
    return OR_OK;
}

/* First FFA shift whose trial period is >= pmax samples (periodogram.hpp:54-57) */
size_t or_ceilshift(size_t rows, size_t cols, double pmax)
{
    return (size_t)ceil(cols * (rows - 1.0) * (1.0 - cols / pmax));
}

static size_t min3(size_t a, size_t b, size_t c)
{
    size_t m = a < b ? a : b;
    return m < c ? m : c;
}

int or_periodogram_length(size_t size, double tsamp, double pmin, double pmax,
                          size_t bmin, size_t bmax, size_t* length)
{
    int st = or_periodogram_check(size, tsamp, pmin, pmax, bmin, bmax);
    if (st) return st;
    const double ds_ini = pmin / (tsamp * bmin);
    const double ds_geo = (bmax + 1.0) / bmin;
    const size_t nds = (size_t)ceil(log(pmax / pmin) / log(ds_geo));
    size_t total = 0;
    for (size_t ids = 0; ids < nds; ++ids) {
        const double f = ds_ini * pow(ds_geo, ids);
        const double tau = f * tsamp;
        const double pmax_samples = pmax / tau;
        const size_t n = or_downsampled_size(size, f);
        const size_t bstop = min3(bmax, n, (size_t)pmax_samples);
        for (size_t bins = bmin; bins <= bstop; ++bins) {
            const size_t rows = n / bins;
            const double pceil = fmin(pmax_samples, bins + 1.0);
            const size_t ce = or_ceilshift(rows, bins, pceil);
            total += rows < ce ? rows : ce;
        }
    }
    *length = total;
    return OR_OK;
}

/* Downsample -> FFA -> S/N for every (ds step, bins) pair, in the
 * reference's plan order (periodogram.hpp:117-201). */
int or_periodogram(const float* data, size_t size, double tsamp, const uint64_t* widths,
                   size_t nw, double pmin, double pmax, size_t bmin, size_t bmax,
                   double* periods, uint32_t* foldbins, float* snr)
{
    int st = or_periodogram_check(size, tsamp, pmin, pmax, bmin, bmax);
    if (st) return st;
    const double ds_ini = pmin / (tsamp * bmin);
    const double ds_geo = (bmax + 1.0) / bmin;
    const size_t nds = (size_t)ceil(log(pmax / pmin) / log(ds_geo));
    const size_t cap = or_downsampled_size(size, ds_ini);
    float* dsbuf = (float*)malloc((cap ? cap : 1) * sizeof(float));
    float* tmp = (float*)malloc((cap ? cap : 1) * sizeof(float));
    float* ffa = (float*)malloc((cap ? cap : 1) * sizeof(float));
    size_t wmax = 1;
    for (size_t i = 0; i < nw; ++i)
        if (widths[i] > wmax) wmax = widths[i];
    float* scratch = (float*)malloc((bmax + wmax + 1) * sizeof(float));

    for (size_t ids = 0; ids < nds; ++ids) {
        const double f = ds_ini * pow(ds_geo, ids);
        const double tau = f * tsamp;
        const double pmax_samples = pmax / tau;
        const size_t n = or_downsampled_size(size, f);
        const float* x = data;
        if (f != 1) {
            or_downsample(data, size, f, dsbuf);
            x = dsbuf;
        }
        const size_t bstop = min3(bmax, n, (size_t)pmax_samples);
        for (size_t bins = bmin; bins <= bstop; ++bins) {
            const size_t rows = n / bins;
            const float stdnoise = (float)sqrt(rows * or_downsampled_variance(size, f));
            const double pceil = fmin(pmax_samples, bins + 1.0);
            const size_t ce = or_ceilshift(rows, bins, pceil);
            const size_t rows_eval = rows < ce ? rows : ce;
            ffa_rec(x, tmp, ffa, rows, bins);
            for (size_t s = 0; s < rows_eval; ++s)
                snr_profile(ffa + s * bins, bins, widths, nw, stdnoise, scratch, snr + s * nw);
            for (size_t s = 0; s < rows_eval; ++s) {
                periods[s] = tau * bins * bins / (bins - s / (rows - 1.0));
                foldbins[s] = (uint32_t)bins;
            }
            snr += rows_eval * nw;
            periods += rows_eval;
            foldbins += rows_eval;
        }
    }
    free(dsbuf);
    free(tmp);
    free(ffa);
    free(scratch);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Float64 recomputation of the periodogram (for error statements only)      */
/* ------------------------------------------------------------------------ */

/* The same search as or_periodogram with every value held in float64: the
 * downsampling weights (imin + 1 - start, end - imax) are not rounded to
 * float32, the FFA sums and the boxcar S/N (h, b, prefix sums) are float64.
 * The plan, the merge tree and its float32 row map are the algorithm's
 * definition and stay as they are.  Output: float64 S/N, plan order. */
static void ffa_rec_f64(const double* in, double* temp, double* out, size_t m, size_t p)
{
    if (m == 1) {
        memcpy(out, in, p * sizeof(double));
        return;
    }
    const size_t mh = m >> 1;
    const size_t mt = m - mh;
    ffa_rec_f64(in, out, temp, mh, p);
    ffa_rec_f64(in + mh * p, out + mh * p, temp + mh * p, mt, p);
    const double* head = temp;
    const double* tail = temp + mh * p;
    const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
    const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
    for (size_t s = 0; s < m; ++s) {
        const size_t h = m == 2 ? 0 : (size_t)(kh * (float)s + 0.5f);
        const size_t t = m == 2 ? 0 : (size_t)(kt * (float)s + 0.5f);
        const size_t sh = (s - t) % p;
        for (size_t c = 0; c < p; ++c)
            out[s * p + c] = head[h * p + c] + tail[t * p + (c + sh) % p];
    }
}

int or_periodogram_f64(const float* data, size_t size, double tsamp, const uint64_t* widths,
                       size_t nw, double pmin, double pmax, size_t bmin, size_t bmax,
                       double* snr)
{
    int st = or_periodogram_check(size, tsamp, pmin, pmax, bmin, bmax);
    if (st) return st;
    const double ds_ini = pmin / (tsamp * bmin);
    const double ds_geo = (bmax + 1.0) / bmin;
    const size_t nds = (size_t)ceil(log(pmax / pmin) / log(ds_geo));
    const size_t cap = size;
    double* x = (double*)malloc(cap * sizeof(double));
    double* tmp = (double*)malloc(cap * sizeof(double));
    double* ffa = (double*)malloc(cap * sizeof(double));
    double* cps = (double*)malloc((2 * bmax + 1) * sizeof(double));
    for (size_t ids = 0; ids < nds; ++ids) {
        const double f = ds_ini * pow(ds_geo, ids);
        const double tau = f * tsamp;
        const double pmax_samples = pmax / tau;
        const size_t n = or_downsampled_size(size, f);
        if (f != 1) {
            for (size_t k = 0; k < n; ++k) {
                const double start = k * f;
                const double end = start + f;
                const size_t imin = (size_t)floor(start);
                const size_t imax = (size_t)fmin(floor(end), size - 1.0);
                double acc = ((imin + 1) - start) * data[imin];
                for (size_t i = imin + 1; i < imax; ++i) acc += data[i];
                acc += (end - imax) * data[imax];
                x[k] = acc;
            }
        } else {
            for (size_t k = 0; k < n; ++k) x[k] = data[k];
        }
        const size_t bstop = min3(bmax, n, (size_t)pmax_samples);
        for (size_t bins = bmin; bins <= bstop; ++bins) {
            const size_t rows = n / bins;
            const double stdnoise = sqrt(rows * or_downsampled_variance(size, f));
            const double pceil = fmin(pmax_samples, bins + 1.0);
            const size_t ce = or_ceilshift(rows, bins, pceil);
            const size_t rows_eval = rows < ce ? rows : ce;
            ffa_rec_f64(x, tmp, ffa, rows, bins);
            for (size_t r = 0; r < rows_eval; ++r) {
                const double* prof = ffa + r * bins;
                cps[0] = 0.0;
                for (size_t j = 0; j < 2 * bins; ++j) cps[j + 1] = cps[j] + prof[j % bins];
                const double total = cps[bins];
                for (size_t iw = 0; iw < nw; ++iw) {
                    const size_t w = widths[iw];
                    const double h = sqrt((bins - w) / (double)(bins * w));
                    const double b = w / (double)(bins - w) * h;
                    double best = -INFINITY;
                    for (size_t i = 0; i < bins; ++i)
                        if (cps[i + w] - cps[i] > best) best = cps[i + w] - cps[i];
                    snr[r * nw + iw] = ((h + b) * best - b * total) / stdnoise;
                }
            }
            snr += rows_eval * nw;
        }
    }
    free(x);
    free(tmp);
    free(ffa);
    free(cps);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Running median: running_median.hpp:13-132                                 */
/* ------------------------------------------------------------------------ */

/* Lomuto partition around data[ipiv], pivot parked at `end` (running_median.hpp:13-35) */
static size_t rm_partition(float* d, size_t start, size_t end, size_t ipiv)
{
    const float pivot = d[ipiv];
    float t = d[ipiv]; d[ipiv] = d[end]; d[end] = t;
    size_t j = start;
    for (size_t i = start; i < end; ++i) {
        if (d[i] < pivot) {
            t = d[i]; d[i] = d[j]; d[j] = t;
            ++j;
        }
    }
    t = d[j]; d[j] = d[end]; d[end] = t;
    return j;
}

/* Quickselect of rank n within [start, end], last element as pivot
 * (running_median.hpp:39-56), written as a loop. */
static float rm_select(float* d, size_t start, size_t end, size_t n)
{
    for (;;) {
        if (start == end)
            return d[start];
        const size_t j = rm_partition(d, start, end, end);
        if (j == n)
            return d[j];
        if (j < n)
            start = j + 1;
        else
            end = j - 1;
    }
}

/* Edge-padded running median: ring buffer of `width` values, median of a
 * copy of the ring after each push (running_median.hpp:59-132). */
int or_running_median(const float* data, size_t size, size_t width, float* out)
{
    if (!(width % 2))
        return fail("width must be an odd number");
    if (!(width < size))
        return fail("width must be < size");
    const size_t half = width / 2;
    float* ring = (float*)malloc(width * sizeof(float));
    float* work = (float*)malloc(width * sizeof(float));
    size_t pos = 0;
    float med = 0.0f;
#define RM_PUSH(v)                                            \
    do {                                                      \
        ring[pos] = (v);                                      \
        pos = (pos + 1 >= width) ? 0 : pos + 1;               \
        memcpy(work, ring, width * sizeof(float));            \
        med = rm_select(work, 0, width - 1, width / 2);       \
    } while (0)
    /* The ring starts zero-filled; the warm-up pushes overwrite all of it. */
    memset(ring, 0, width * sizeof(float));
    for (size_t i = 0; i < half + 1; ++i)
        RM_PUSH(data[0]);
    for (size_t i = 1; i < half; ++i)
        RM_PUSH(data[i]);
    for (size_t i = half; i < size; ++i) {
        RM_PUSH(data[i]);
        out[i - half] = med;
    }
    for (size_t i = 0; i < half; ++i) {
        RM_PUSH(data[size - 1]);
        out[size - half + i] = med;
    }
#undef RM_PUSH
    free(ring);
    free(work);
    return OR_OK;
}
