// Per-series kernels for gfx950: fractional downsampling (single factor and
// all factors of a search in one pass), running median, the block means and
// interpolation of the dereddening, normalisation, and the small helpers of
// the reference's binding (rollback, fused_rollback_add, circular prefix sum).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "exact_f64.h"
#include "ffa_device.h"
#include "series_kernels.h"

namespace rt {

// ---------------------------------------------------------------------------
// Downsampling (downsample.hpp:44-82)
// ---------------------------------------------------------------------------

// One thread per output sample, inputs read straight from HBM.  Used for
// factors too large for the LDS-staged kernel.
__global__ void downsample_direct_kernel(const float* __restrict__ in, int64_t N, uint64_t mf,
                                         int ef, int64_t n, float* __restrict__ out)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const DsIndex ix = ds_index_fast(k, mf, ef, N);
        out[k] = ds_accumulate(in, ix);
    }
}

// Folding (folding.py:21-85): rows x bins samples scaled in place by a
// float32 factor (`folded *= (m * factor) ** -0.5`).
__global__ void fold_scale_kernel(float* __restrict__ x, int64_t count, float scale)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= scale;
}

// Sum over the rows of a rows x bins array, row after row (numpy's reduction
// over the outer axis of a C-ordered array adds the rows in order).
__global__ void fold_sum_rows_kernel(const float* __restrict__ x, int64_t rows, int64_t bins,
                                     float* __restrict__ out)
{
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < bins;
         b += (int64_t)gridDim.x * blockDim.x) {
        float acc = x[b];
        for (int64_t i = 1; i < rows; ++i) acc += x[(size_t)i * bins + b];
        out[b] = acc;
    }
}

// Downsampling of every column of a rows x bins array by f
// (folding.py:6-18): output row k of column b accumulates input rows
// [imin, imax] of that column exactly as the one-series downsampling does.
__global__ void downsample_vertical_kernel(const float* __restrict__ x, int64_t rows,
                                           int64_t bins, uint64_t mf, int ef, int64_t n,
                                           float* __restrict__ out)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * bins;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / bins;
        const int64_t b = e - k * bins;
        const DsIndex ix = ds_index_fast(k, mf, ef, rows);
        const float* col = x + b;
        float acc = ix.wmin * col[(size_t)ix.imin * bins];
        for (int64_t i = ix.imin + 1; i < ix.imax; ++i) acc += col[(size_t)i * bins];
        out[e] = __fmaf_rn(ix.wmax, col[(size_t)ix.imax * bins], acc);
    }
}

// 8-bit samples of a staged file to float32 (kind 1: unsigned, 2: signed).
__global__ void convert_samples_kernel(const uint8_t* __restrict__ raw, int64_t n, int kind,
                                       float* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = kind == 2 ? (float)(int8_t)raw[i] : (float)raw[i];
}

// First output sample whose imin is >= c.
__device__ __forceinline__ int64_t ds_first_at(int64_t c, const DsStepDev& st)
{
    if (c <= 0) return 0;
    int64_t k = (int64_t)((float)c * st.inv_f);  // float estimate, fixed below
    if (k < 0) k = 0;
    if (k > st.n) k = st.n;
    while (k > 0 && ds_imin(k - 1, st.mf, st.ef) >= c) --k;
    while (k < st.n && ds_imin(k, st.mf, st.ef) < c) ++k;
    return k;
}

// The reference's sum of staged samples in[ia..ib] (downsample.hpp:73-78):
// wmin*in[ia], the interior samples left to right, then the last one fused.
// The first NI interior samples are loaded together (a factor f has at most
// floor(f) + 1 of them; more than NI take the loop at the end).
// Staged sample i sits at LDS word dsp(i): one pad word per 32 samples, so
// that the lanes of a wavefront reading samples about f apart (consecutive
// outputs of a factor f) fall on different banks.
__host__ __device__ inline int dsp(int i) { return i + (i >> 5); }
// LDS words of an array of n padded samples.
__host__ __device__ inline int ds_words(int n) { return n + (n >> 5) + 1; }

template <int NI>
__device__ __forceinline__ float ds_direct(const float* in, int ia, int ib, float wmin, float wmax)
{
    const int n = ib - ia - 1;  // interior samples (-1: ia == ib)
    const int last = n > 1 ? n - 1 : 0;
    float t[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) t[j] = in[dsp(ia + 1 + min(j, last))];
    float acc = wmin * in[dsp(ia)];
#pragma unroll
    for (int j = 0; j < NI; ++j)
        if (j < n) acc += t[j];
    for (int i = ia + 1 + NI; i < ib; ++i) acc += in[dsp(i)];
    return __fmaf_rn(wmax, in[dsp(ib)], acc);
}

// All downsampling steps of a search in one pass over the raw series: every
// workgroup stages one chunk of the input (plus a halo of the largest factor)
// in LDS and produces, for every step, the outputs whose first input sample
// lies in the chunk.  The staged samples also get an exclusive float-float
// prefix sum, so that an output of a large factor costs its two edge terms
// and one prefix difference (the interior sum to within about one rounding)
// instead of a chain of f dependent adds; factors up to kDsDirectMax sum
// their few interior samples directly, left to right.  Thread t takes the
// flat outputs t, t + blockDim, ... of the chunk in step order, following
// them with a step index that only moves forward.
constexpr float kDsDirectMax = 12.0f;

__global__ __launch_bounds__(kDsThreads) void downsample_packed_kernel(
    const DsSeriesDev* __restrict__ series, int nseries, const DsStepDev* __restrict__ steps,
    int chunk, int halo)
{
    extern __shared__ float s_mem[];
    __shared__ int64_t s_kbeg[kDsMaxSteps];
    __shared__ int s_cnt_prefix[kDsMaxSteps + 1];
    __shared__ float s_th[kDsThreads], s_tl[kDsThreads];
    // per step: the factor as a double, its reciprocal and the output row
    __shared__ double s_f[kDsMaxSteps];
    __shared__ float s_invf[kDsMaxSteps];
    __shared__ float* s_out[kDsMaxSteps];
    const int cap = chunk + halo;
    float* s_in = s_mem;             // staged samples
    float* s_ch = s_in + ds_words(cap + 1);  // exclusive prefix, high parts (cap + 1)
    float* s_cl = s_ch + ds_words(cap + 1);  // low parts

    // find the series of this chunk: the proportional guess (exact when the
    // series have equal chunk counts), else a binary search
    const int b = blockIdx.x;
    int lo = (int)min((int64_t)b * nseries / (int64_t)gridDim.x, (int64_t)nseries - 1);
    if (!(series[lo].chunk_offset <= b &&
          (lo == nseries - 1 || b < series[lo + 1].chunk_offset))) {
        int hi = nseries - 1;
        lo = 0;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (series[mid].chunk_offset <= b) lo = mid;
            else hi = mid - 1;
        }
    }
    const DsSeriesDev sr = series[lo];
    const int64_t c0 = (int64_t)(b - sr.chunk_offset) * chunk;
    const int64_t c1 = min(c0 + chunk, sr.N);
    const int L = (int)(min(c0 + chunk + halo, sr.N) - c0);
    const int tid = threadIdx.x;
    {
        // every load of the chunk issued before the first LDS store
        constexpr int R = kDsStageLoads;
        float v[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int j = tid + u * kDsThreads;
            if (j < L) v[u] = sr.x[c0 + j];
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int j = tid + u * kDsThreads;
            if (j < L) s_in[dsp(j)] = v[u];
        }
        for (int j = tid + R * kDsThreads; j < L; j += kDsThreads) s_in[dsp(j)] = sr.x[c0 + j];
    }

    const DsStepDev* st = steps + sr.step0;
    for (int s = tid; s < sr.nsteps; s += blockDim.x) {
        const int64_t kb = ds_first_at(c0, st[s]);
        const int64_t ke = (c1 >= sr.N) ? st[s].n : ds_first_at(c1, st[s]);
        s_kbeg[s] = kb;
        s_cnt_prefix[s + 1] = (int)(ke - kb);
        s_f[s] = ds_factor(st[s].mf, st[s].ef);
        s_invf[s] = st[s].inv_f;
        s_out[s] = sr.out + st[s].out_offset;
    }
    __syncthreads();
    // float-float running sum of this thread's segment of the staged samples
    const int seg = (L + (int)blockDim.x - 1) / (int)blockDim.x;
    const int j0 = min(L, tid * seg), j1 = min(L, j0 + seg);
    float sh = 0.0f, sl = 0.0f;
    for (int j = j0; j < j1; ++j) {
        float t, e;
        two_sum(sh, s_in[dsp(j)], t, e);
        sh = t;
        sl += e;
    }
    s_th[tid] = sh;
    s_tl[tid] = sl;
    if (tid == 0) {
        s_cnt_prefix[0] = 0;
        for (int s = 0; s < sr.nsteps; ++s)
            s_cnt_prefix[s + 1] += s_cnt_prefix[s];
    }
    __syncthreads();
    // exclusive scan of the thread totals by the first wavefront: lane l
    // takes totals Q*l .. Q*l + Q - 1
    constexpr int Q = kDsThreads / kWave;
    static_assert(Q * kWave == kDsThreads, "threads must be whole wavefronts");
    if (tid < kWave) {
        float ph[Q], pl[Q];
        float ah = 0.0f, al = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            ph[q] = ah;
            pl[q] = al;
            ff_add(ah, al, s_th[Q * tid + q], s_tl[Q * tid + q], ah, al);
        }
        float ih = ah, il = al;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const float oh = __shfl_up(ih, d, kWave);
            const float ol = __shfl_up(il, d, kWave);
            if (tid >= d) ff_add(oh, ol, ih, il, ih, il);
        }
        float eh = __shfl_up(ih, 1, kWave);
        float el = __shfl_up(il, 1, kWave);
        if (tid == 0) {
            eh = 0.0f;
            el = 0.0f;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float rh, rl;
            ff_add(eh, el, ph[q], pl[q], rh, rl);
            s_th[Q * tid + q] = rh;
            s_tl[Q * tid + q] = rl;
        }
    }
    __syncthreads();
    // the segment rescanned from its offset: C[j] = sum of s_in[0 .. j)
    sh = s_th[tid];
    sl = s_tl[tid];
    for (int j = j0; j < j1; ++j) {
        s_ch[dsp(j)] = sh + sl;
        s_cl[dsp(j)] = sl - ((sh + sl) - sh);
        float t, e;
        two_sum(sh, s_in[dsp(j)], t, e);
        sh = t;
        sl += e;
    }
    if (j1 == L && j0 < j1) {
        s_ch[dsp(L)] = sh + sl;
        s_cl[dsp(L)] = sl - ((sh + sl) - sh);
    }
    __syncthreads();

    const int total = s_cnt_prefix[sr.nsteps];
    // series below 2^31 samples: the index math in native float64
    const bool f64 = sr.N < ((int64_t)1 << 31);
    const int n32 = (int)min(sr.N, (int64_t)INT32_MAX);
    // the step of the current output, its descriptor in registers, moved on
    // when the flat output index crosses into a later step
    int a = 0, a_first = 0, a_end = s_cnt_prefix[1];
    int64_t kb = s_kbeg[0];
    double f = s_f[0];
    float invf = s_invf[0];
    float* orow = s_out[0];
    for (int e = tid; e < total; e += blockDim.x) {
        if (e >= a_end) {
            do {
                ++a;
            } while (s_cnt_prefix[a + 1] <= e);
            a_first = s_cnt_prefix[a];
            a_end = s_cnt_prefix[a + 1];
            kb = s_kbeg[a];
            f = s_f[a];
            invf = s_invf[a];
            orow = s_out[a];
        }
        const int64_t k = kb + (e - a_first);
        const DsIndex ix = f64 ? ds_index_f64((int)k, f, n32)
                               : ds_index_fast(k, st[a].mf, st[a].ef, sr.N);
        const int ia = (int)(ix.imin - c0), ib = (int)(ix.imax - c0);
        float v;
        if (invf >= 1.0f / 4.0f) {
            v = ds_direct<5>(s_in, ia, ib, ix.wmin, ix.wmax);
        } else if (invf >= 1.0f / 8.0f) {
            v = ds_direct<9>(s_in, ia, ib, ix.wmin, ix.wmax);
        } else if (invf >= 1.0f / kDsDirectMax) {
            v = ds_direct<13>(s_in, ia, ib, ix.wmin, ix.wmax);
        } else {
            float inner = 0.0f;
            if (ib > ia + 1)
                inner = (s_ch[dsp(ib)] - s_ch[dsp(ia + 1)]) + (s_cl[dsp(ib)] - s_cl[dsp(ia + 1)]);
            v = __fmaf_rn(ix.wmax, s_in[dsp(ib)], ix.wmin * s_in[dsp(ia)] + inner);
        }
        orow[k] = v;
    }
}

// ---------------------------------------------------------------------------
// Running median (running_median.hpp:100-132): exact order statistic of the
// edge-padded window (pure comparisons, bit-exact).
// ---------------------------------------------------------------------------

// LDS floats of running_median_block's sorted path for `threads` outputs.
__host__ __device__ inline size_t rmed_lds_floats(int threads, int width)
{
    return (size_t)3 * (threads + width - 1);
}

// Median of the edge-padded window of `width` (odd) samples centred on
// output o0 + threadIdx.x.
//  sorted (use_lds): the workgroup's union window x[o0 - half .. o0 +
//    blockDim.x + half) (clamped) goes to LDS and every sample gets its
//    stable rank (samples below it, plus equal ones before it) by one pass of
//    comparisons; then each output walks the union in sorted order, counting
//    the samples of its own window, and takes the one that completes half + 1.
//    About 2 (blockDim + width) comparisons per output.
//  counting (otherwise): each output tests the samples of its window in turn,
//    counting those below and not above each candidate (about width^2 / 2
//    comparisons), reading x directly.
__device__ __forceinline__ void running_median_block(const float* __restrict__ x, int64_t N,
                                                     int width, int use_lds, int64_t o0,
                                                     float* s_win, float* __restrict__ out)
{
    const int half = width / 2;
    const int span = blockDim.x + width - 1;
    const int t = threadIdx.x;
    const int64_t i = o0 + t;
    if (use_lds) {
        float* s_val = s_win;                                  // window samples
        float* s_sv = s_win + span;                            // sorted samples
        int* s_si = reinterpret_cast<int*>(s_win + 2 * span);  // their window positions
        for (int j = t; j < span; j += blockDim.x) {
            int64_t src = o0 + j - half;
            src = src < 0 ? 0 : (src > N - 1 ? N - 1 : src);
            s_val[j] = x[src];
        }
        __syncthreads();
        for (int u = t; u < span; u += blockDim.x) {
            const float v = s_val[u];
            int r = 0;
            for (int q = 0; q < span; ++q) {
                const float y = s_val[q];
                r += (y < v) || (y == v && q < u);
            }
            s_sv[r] = v;
            s_si[r] = u;
        }
        __syncthreads();
        if (i >= N) return;
        float med = 0.0f;
        int cnt = 0;
        bool found = false;
        for (int r = 0; r < span; ++r) {
            const bool in = (unsigned)(s_si[r] - t) < (unsigned)width;
            cnt += in ? 1 : 0;
            if (!found && in && cnt == half + 1) {
                med = s_sv[r];
                found = true;
            }
            if (__all(found)) break;
        }
        out[i] = med;
        return;
    }
    if (i >= N) return;
    float med = 0.0f;
    for (int c = 0; c < width; ++c) {
        int64_t src = i + c - half;
        src = src < 0 ? 0 : (src > N - 1 ? N - 1 : src);
        const float cand = x[src];
        int lt = 0, le = 0;
        for (int j = 0; j < width; ++j) {
            int64_t sj = i + j - half;
            sj = sj < 0 ? 0 : (sj > N - 1 ? N - 1 : sj);
            const float v = x[sj];
            lt += (v < cand);
            le += (v <= cand);
        }
        if (lt <= half && half < le) {
            med = cand;
            break;
        }
    }
    out[i] = med;
}

__global__ __launch_bounds__(256) void running_median_kernel(const float* __restrict__ x,
                                                             int64_t N, int width,
                                                             int use_lds,
                                                             float* __restrict__ out)
{
    extern __shared__ float s_win[];
    running_median_block(x, N, width, use_lds, (int64_t)blockIdx.x * blockDim.x, s_win, out);
}

// ---------------------------------------------------------------------------
// Dereddening (running_medians.py:40-85, time_series.py:95-124)
// ---------------------------------------------------------------------------

// numpy's float32 pairwise summation (8 accumulators on blocks of <= 128,
// halving split otherwise) of a[0..n).
__device__ __forceinline__ float pairwise_leaf(const float* a, int64_t n)
{
    if (n < 8) {
        float res = 0.0f;
        for (int64_t i = 0; i < n; ++i) res += a[i];
        return res;
    }
    float r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3];
    float r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8) {
        r0 += a[i + 0]; r1 += a[i + 1]; r2 += a[i + 2]; r3 += a[i + 3];
        r4 += a[i + 4]; r5 += a[i + 5]; r6 += a[i + 6]; r7 += a[i + 7];
    }
    float res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; ++i) res += a[i];
    return res;
}

__device__ float pairwise_sum(const float* a, int64_t n)
{
    // explicit stack replaces numpy's recursion; combines results in the
    // same (left + right) order
    struct Frame {
        const float* a;
        int64_t n;
        float left;
        int state;
    };
    Frame st[40];
    int sp = 0;
    st[0].a = a;
    st[0].n = n;
    st[0].state = 0;
    float ret = 0.0f;
    for (;;) {
        Frame& f = st[sp];
        if (f.n <= 128) {
            ret = pairwise_leaf(f.a, f.n);
            if (sp == 0) return ret;
            --sp;
            continue;  // deliver ret to the parent
        }
        int64_t n2 = f.n / 2;
        n2 -= n2 % 8;
        if (f.state == 0) {
            f.state = 1;
            Frame& c = st[++sp];
            c.a = f.a;
            c.n = n2;
            c.state = 0;
        } else if (f.state == 1) {
            f.left = ret;
            f.state = 2;
            Frame& c = st[++sp];
            c.a = f.a + n2;
            c.n = f.n - n2;
            c.state = 0;
        } else {
            ret = f.left + ret;
            if (sp == 0) return ret;
            --sp;
        }
    }
}

// Block means of `factor` samples: numpy float32 mean = (0 + pairwise sum)/factor.
__device__ __forceinline__ void scrunch_mean_range(const float* __restrict__ x, int64_t nblocks,
                                                   int64_t factor, float* __restrict__ out,
                                                   int64_t first, int64_t stride)
{
    for (int64_t b = first; b < nblocks; b += stride) {
        const float s = 0.0f + pairwise_sum(x + b * factor, factor);
        out[b] = s / (float)factor;
    }
}

__global__ void scrunch_mean_kernel(const float* __restrict__ x, int64_t nblocks, int64_t factor,
                                    float* __restrict__ out)
{
    scrunch_mean_range(x, nblocks, factor, out, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                       (int64_t)gridDim.x * blockDim.x);
}

// Samples of the elementwise preparation kernels are taken four at a time
// (group g = samples 4g .. 4g + 3): one 16-byte access per lane when both
// buffers are 16-byte aligned, scalar accesses otherwise, the same arithmetic
// either way.
__device__ __forceinline__ int load_group(const float* __restrict__ x, int64_t i0, int64_t N,
                                          bool vec, float (&v)[4])
{
    const int cnt = (int)min((int64_t)4, N - i0);
    if (vec && cnt == 4) {
        const float4 t = *reinterpret_cast<const float4*>(x + i0);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
        v[3] = t.w;
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = u < cnt ? x[i0 + u] : 0.0f;
    }
    return cnt;
}

__device__ __forceinline__ void store_group(float* __restrict__ out, int64_t i0, int cnt, bool vec,
                                            const float (&v)[4])
{
    if (vec && cnt == 4) {
        *reinterpret_cast<float4*>(out + i0) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < cnt) out[i0 + u] = v[u];
    }
}

__device__ __forceinline__ bool aligned16(const void* a, const void* b)
{
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

// numpy's linear interpolation of the low-resolution running median fp at
// sample i, as a float-float (vh, vl): num = 2i - sf + 1 and j the last node
// with 2*x_j = 2*j*sf + sf - 1 <= 2i (-1 before the first), clamped to the
// end values, exact node values on the nodes.
__device__ __forceinline__ void rmed_interp(const float* __restrict__ fp, int64_t nl, int64_t sf,
                                            int64_t num, int64_t j, float& vh, float& vl)
{
    if (j < 0) {
        vh = fp[0];
        vl = 0.0f;
    } else if (j >= nl - 1) {
        vh = fp[nl - 1];
        vl = 0.0f;
    } else {
        const int64_t t2 = num - 2 * j * sf;  // 2*(i - x_j), in [0, 2*sf)
        if (t2 == 0) {
            vh = fp[j];
            vl = 0.0f;
        } else {
            // (fp[j+1] - fp[j]) * (t2/2) / sf + fp[j]
            float dh, dl;
            two_sum(fp[j + 1], -fp[j], dh, dl);
            const float t = 0.5f * (float)t2;  // exact (t2 < 2^24)
            const float ph = dh * t;
            const float pl = __fmaf_rn(dh, t, -ph) + dl * t;
            const float fsf = (float)sf;
            const float qh = ph / fsf;
            const float r = __fmaf_rn(-qh, fsf, ph) + pl;
            const float ql = r / fsf;
            float sh, se;
            two_sum(fp[j], qh, sh, se);
            vh = sh;
            vl = se + ql;
        }
    }
}

// out[i] = float32(x[i] - interp(i)) (or float32(interp(i)) when x is null),
// where interp is numpy's linear interpolation of the low-resolution running
// median at centres j*sf + (sf-1)/2 (np.interp semantics), evaluated in
// float-float arithmetic and rounded once.  Groups of four samples from
// `first`, `stride` groups apart; one integer division per group (the node
// index moves by at most one per sample).
__device__ __forceinline__ void deredden_subtract_range(const float* __restrict__ x, int64_t N,
                                                        const float* __restrict__ fp, int64_t nl,
                                                        int64_t sf, float* __restrict__ out,
                                                        int64_t first, int64_t stride)
{
    const bool vec = aligned16(x, out);
    const int64_t two_sf = 2 * sf;
    for (int64_t g = first; 4 * g < N; g += stride) {
        const int64_t i0 = 4 * g;
        float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int cnt = (int)min((int64_t)4, N - i0);
        if (x != nullptr) cnt = load_group(x, i0, N, vec, xv);
        int64_t num = 2 * i0 - sf + 1;
        int64_t j = (num >= 0) ? num / two_sf : -1;
        float r[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u > 0) {
                num += 2;
                if (num >= two_sf * (j + 1)) ++j;
            }
            float vh, vl;
            rmed_interp(fp, nl, sf, num, j, vh, vl);
            if (x == nullptr) {
                r[u] = vh + vl;  // the interpolated running median alone
            } else {
                float d, e;
                two_sum(xv[u], -vh, d, e);
                e -= vl;
                r[u] = d + e;
            }
        }
        store_group(out, i0, cnt, vec, r);
    }
}

__global__ void deredden_subtract_kernel(const float* __restrict__ x, int64_t N,
                                         const float* __restrict__ fp, int64_t nl, int64_t sf,
                                         float* __restrict__ out)
{
    deredden_subtract_range(x, N, fp, nl, sf, out, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                            (int64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Normalisation (time_series.py:67-93): float-float sums, fixed reduction
// order, one rounding of (x - mean)/std.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void block_ff_reduce(float& h, float& l, float* s_h, float* s_l)
{
    // wavefront reduce then across wavefronts, fixed order
    for (int d = 32; d >= 1; d >>= 1) {
        const float oh = __shfl_down(h, d, kWave);
        const float ol = __shfl_down(l, d, kWave);
        ff_add(h, l, oh, ol, h, l);
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0) {
        s_h[wave] = h;
        s_l[wave] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float rh = s_h[0], rl = s_l[0];
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w)
            ff_add(rh, rl, s_h[w], s_l[w], rh, rl);
        h = rh;
        l = rl;
    }
}

// mode 0: partial sums of x; mode 1: partial sums of (x - mean)^2; part
// `part` of `nparts` (256 threads each): thread t of part p takes the groups
// of four samples p*256 + t, (nparts + p)*256 + t, ... in sample order
__device__ __forceinline__ void moments_partial(const float* __restrict__ x, int64_t N, int mode,
                                                const float* __restrict__ mean,
                                                float* __restrict__ partial, int part, int nparts,
                                                float* s_h, float* s_l)
{
    float h = 0.0f, l = 0.0f;
    const float mh = mode ? mean[0] : 0.0f;
    const float ml = mode ? mean[1] : 0.0f;
    const bool vec = aligned16(x, x);
    for (int64_t g = (int64_t)part * blockDim.x + threadIdx.x; 4 * g < N;
         g += (int64_t)nparts * blockDim.x) {
        float v[4];
        const int cnt = load_group(x, 4 * g, N, vec, v);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= cnt) break;
            if (mode == 0) {
                ff_add(h, l, v[u], 0.0f, h, l);
            } else {
                float dh, dl;
                two_sum(v[u], -mh, dh, dl);
                dl -= ml;
                const float sh = dh * dh;
                const float sl = __fmaf_rn(dh, dh, -sh) + 2.0f * dh * dl;
                ff_add(h, l, sh, sl, h, l);
            }
        }
    }
    block_ff_reduce(h, l, s_h, s_l);
    if (threadIdx.x == 0) {
        partial[2 * part] = h;
        partial[2 * part + 1] = l;
    }
}

__global__ __launch_bounds__(256) void moments_partial_kernel(const float* __restrict__ x,
                                                              int64_t N, int mode,
                                                              const float* __restrict__ mean,
                                                              float* __restrict__ partial)
{
    __shared__ float s_h[4], s_l[4];
    moments_partial(x, N, mode, mean, partial, blockIdx.x, gridDim.x, s_h, s_l);
}

// Reduces the partials; mode 0 writes mean = sum/N, mode 1 writes
// std = sqrt(sum/N), both as float-float pairs.
__device__ __forceinline__ void moments_final(const float* __restrict__ partial, int nparts,
                                              int64_t N, int mode, float* __restrict__ result,
                                              float* s_h, float* s_l)
{
    float h = 0.0f, l = 0.0f;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        ff_add(h, l, partial[2 * i], partial[2 * i + 1], h, l);
    block_ff_reduce(h, l, s_h, s_l);
    if (threadIdx.x == 0) {
        // divide the float-float sum by N (exact in float for N < 2^24 only:
        // use a float-float divisor)
        const float nh = (float)N;
        const float nl = (float)(N - (int64_t)nh);
        float qh = h / nh;
        // remainder r = (h + l) - qh*(nh + nl)
        const float p = qh * nh;
        const float pe = __fmaf_rn(qh, nh, -p);
        float r = ((h - p) - pe) + l - qh * nl;
        float ql = r / nh;
        float mh, ml;
        mh = qh + ql;
        ml = ql - (mh - qh);
        if (mode == 1) {
            // sqrt of a float-float: one Newton step from the float root
            const float s = sqrtf(mh);
            const float ss = s * s;
            const float sse = __fmaf_rn(s, s, -ss);
            const float corr = ((mh - ss) - sse + ml) / (2.0f * s);
            mh = s + corr;
            ml = corr - (mh - s);
        }
        result[0] = mh;
        result[1] = ml;
    }
}

__global__ __launch_bounds__(256) void moments_final_kernel(const float* __restrict__ partial,
                                                            int nparts, int64_t N, int mode,
                                                            float* __restrict__ result)
{
    __shared__ float s_h[4], s_l[4];
    moments_final(partial, nparts, N, mode, result, s_h, s_l);
}

// out = float32((x - mean) / std) with float-float mean and std, groups of
// four samples from `first`, `stride` groups apart
__device__ __forceinline__ void normalise_apply_range(const float* __restrict__ x, int64_t N,
                                                      const float* __restrict__ mean,
                                                      const float* __restrict__ stdv,
                                                      float* __restrict__ out, int64_t first,
                                                      int64_t stride)
{
    const float mh = mean[0], ml = mean[1];
    const float sh = stdv[0], sl = stdv[1];
    const bool vec = aligned16(x, out);
    for (int64_t g = first; 4 * g < N; g += stride) {
        const int64_t i0 = 4 * g;
        float v[4];
        const int cnt = load_group(x, i0, N, vec, v);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float dh, dl;
            two_sum(v[u], -mh, dh, dl);
            dl -= ml;
            // (dh + dl) / (sh + sl)
            const float q = dh / sh;
            const float p = q * sh;
            const float pe = __fmaf_rn(q, sh, -p);
            const float r = ((dh - p) - pe) + dl - q * sl;
            v[u] = q + r / sh;
        }
        store_group(out, i0, cnt, vec, v);
    }
}

__global__ void normalise_apply_kernel(const float* __restrict__ x, int64_t N,
                                       const float* __restrict__ mean,
                                       const float* __restrict__ stdv, float* __restrict__ out)
{
    normalise_apply_range(x, N, mean, stdv, out, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                          (int64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Batched preparation: the kernels above over many series at once, series
// = blockIdx.y, same per-series arithmetic and partition as the one-series
// launches (bit-identical results).
// ---------------------------------------------------------------------------

__global__ void prep_scrunch_kernel(const PrepSeriesDev* __restrict__ ps)
{
    const PrepSeriesDev p = ps[blockIdx.y];
    if (p.sf <= 1) return;
    scrunch_mean_range(p.x, p.nl, p.sf, p.lores, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                       (int64_t)gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(256) void prep_rmed_kernel(const PrepSeriesDev* __restrict__ ps,
                                                        int use_lds)
{
    extern __shared__ float s_win[];
    const PrepSeriesDev p = ps[blockIdx.y];
    const float* x = p.sf > 1 ? p.lores : p.x;
    const int64_t n = p.sf > 1 ? p.nl : p.N;
    for (int64_t o0 = (int64_t)blockIdx.x * blockDim.x; o0 < n;
         o0 += (int64_t)gridDim.x * blockDim.x) {
        running_median_block(x, n, p.width, use_lds, o0, s_win, p.rmed);
        __syncthreads();
    }
}

__global__ void prep_subtract_kernel(const PrepSeriesDev* __restrict__ ps)
{
    const PrepSeriesDev p = ps[blockIdx.y];
    deredden_subtract_range(p.x, p.N, p.rmed, p.sf > 1 ? p.nl : p.N, p.sf, p.out,
                            (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                            (int64_t)gridDim.x * blockDim.x);
}

// normalisation input: the dereddened series (out) or the raw one
__device__ __forceinline__ const float* prep_norm_src(const PrepSeriesDev& p)
{
    return p.dereddened ? p.out : p.x;
}

__global__ __launch_bounds__(256) void prep_moments_partial_kernel(
    const PrepSeriesDev* __restrict__ ps, int mode)
{
    __shared__ float s_h[4], s_l[4];
    const PrepSeriesDev p = ps[blockIdx.y];
    moments_partial(prep_norm_src(p), p.N, mode, p.stats + 2 * kNormaliseParts, p.stats,
                    blockIdx.x, gridDim.x, s_h, s_l);
}

__global__ __launch_bounds__(256) void prep_moments_final_kernel(
    const PrepSeriesDev* __restrict__ ps, int mode)
{
    __shared__ float s_h[4], s_l[4];
    const PrepSeriesDev p = ps[blockIdx.y];
    moments_final(p.stats, kNormaliseParts, p.N, mode, p.stats + 2 * kNormaliseParts + 2 * mode,
                  s_h, s_l);
}

__global__ void prep_normalise_apply_kernel(const PrepSeriesDev* __restrict__ ps, int grid_x)
{
    const PrepSeriesDev p = ps[blockIdx.y];
    normalise_apply_range(prep_norm_src(p), p.N, p.stats + 2 * kNormaliseParts,
                          p.stats + 2 * kNormaliseParts + 2, p.out,
                          (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                          (int64_t)grid_x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Binding helpers (kernels.hpp:19-38, 73-101)
// ---------------------------------------------------------------------------

__global__ void rollback_kernel(const float* __restrict__ x, int64_t size, int64_t shift,
                                float* __restrict__ out)
{
    const int64_t p = shift % size;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = i + p;
        if (j >= size) j -= size;
        out[i] = x[j];
    }
}

__global__ void fused_rollback_add_kernel(const float* __restrict__ x,
                                          const float* __restrict__ y, int64_t size,
                                          int64_t shift, float* __restrict__ z)
{
    const int64_t p = shift % size;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = i + p;
        if (j >= size) j -= size;
        z[i] = x[i] + y[j];
    }
}

// Circular prefix sum over nsum positions: float-float scan of the first
// wrap (rounded to float32), then the float32 wraps out[j] + i*sum.
__global__ __launch_bounds__(64) void circular_prefix_sum_kernel(const float* __restrict__ x,
                                                                 int64_t size, int64_t nsum,
                                                                 float* __restrict__ out)
{
    const int lane = threadIdx.x;
    const int64_t first = size < nsum ? size : nsum;
    snr_prefix_long(x, out, (int)first, lane);
    __syncthreads();
    if (nsum <= size) return;
    const float total = out[size - 1];
    const int64_t wraps = nsum / size;
    const int64_t rest = nsum % size;
    for (int64_t i = 1; i < wraps; ++i) {
        const float a = (float)i * total;
        for (int64_t j = lane; j < size; j += kWave)
            out[i * size + j] = out[j] + a;
    }
    const float a = (float)wraps * total;
    for (int64_t j = lane; j < rest; j += kWave)
        out[wraps * size + j] = out[j] + a;
}

__global__ void fill_zero_kernel(float* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}


// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

void split_double(double f, uint64_t* mf, int* ef)
{
    int e = 0;
    const double m = frexp(f, &e);  // f = m * 2^e, 0.5 <= m < 1
    *mf = (uint64_t)ldexp(m, 53);
    *ef = e - 53;
}

static inline int grid_for(int64_t n, int threads, int cap = 65536)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

hipError_t launch_downsample_direct(const float* in, int64_t N, double f, int64_t n, float* out,
                                    hipStream_t stream)
{
    uint64_t mf;
    int ef;
    split_double(f, &mf, &ef);
    hipLaunchKernelGGL(downsample_direct_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, in,
                       N, mf, ef, n, out);
    return hipGetLastError();
}

hipError_t launch_fold_scale(float* x, int64_t count, float scale, hipStream_t stream)
{
    hipLaunchKernelGGL(fold_scale_kernel, dim3(grid_for(count, 256)), dim3(256), 0, stream, x,
                       count, scale);
    return hipGetLastError();
}

hipError_t launch_fold_sum_rows(const float* x, int64_t rows, int64_t bins, float* out,
                                hipStream_t stream)
{
    hipLaunchKernelGGL(fold_sum_rows_kernel, dim3(grid_for(bins, 64)), dim3(64), 0, stream, x,
                       rows, bins, out);
    return hipGetLastError();
}

hipError_t launch_downsample_vertical(const float* x, int64_t rows, int64_t bins, double f,
                                      int64_t n, float* out, hipStream_t stream)
{
    uint64_t mf;
    int ef;
    split_double(f, &mf, &ef);
    hipLaunchKernelGGL(downsample_vertical_kernel, dim3(grid_for(n * bins, 256)), dim3(256), 0,
                       stream, x, rows, bins, mf, ef, n, out);
    return hipGetLastError();
}

hipError_t launch_convert_samples(const uint8_t* raw, int64_t n, int kind, float* out,
                                  hipStream_t stream)
{
    hipLaunchKernelGGL(convert_samples_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, raw,
                       n, kind, out);
    return hipGetLastError();
}

hipError_t launch_downsample_packed(const DsSeriesDev* d_series, int nseries,
                                    const DsStepDev* d_steps, int total_chunks, int chunk,
                                    int halo, hipStream_t stream)
{
    // staged samples and the two prefix arrays
    const size_t lds = (size_t)3 * ds_words(chunk + halo + 1) * sizeof(float);
    static size_t configured = 0;
    if (lds > 48 * 1024 && lds > configured) {
        const hipError_t e = hipFuncSetAttribute((const void*)downsample_packed_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
        configured = lds;
    }
    hipLaunchKernelGGL(downsample_packed_kernel, dim3(total_chunks), dim3(kDsThreads), lds, stream,
                       d_series, nseries, d_steps, chunk, halo);
    return hipGetLastError();
}

hipError_t launch_running_median(const float* x, int64_t N, int width, float* out,
                                 hipStream_t stream)
{
    const int threads = 256;
    const size_t lds = rmed_lds_floats(threads, width) * sizeof(float);
    const int use_lds = lds <= 48 * 1024;
    hipLaunchKernelGGL(running_median_kernel, dim3((unsigned)((N + threads - 1) / threads)),
                       dim3(threads), use_lds ? lds : 0, stream, x, N, width, use_lds, out);
    return hipGetLastError();
}

hipError_t launch_scrunch_mean(const float* x, int64_t nblocks, int64_t factor, float* out,
                               hipStream_t stream)
{
    hipLaunchKernelGGL(scrunch_mean_kernel, dim3(grid_for(nblocks, 64)), dim3(64), 0, stream, x,
                       nblocks, factor, out);
    return hipGetLastError();
}

hipError_t launch_deredden_subtract(const float* x, int64_t N, const float* fp, int64_t nl,
                                    int64_t sf, float* out, hipStream_t stream)
{
    hipLaunchKernelGGL(deredden_subtract_kernel, dim3(grid_for((N + 3) / 4, 256)), dim3(256), 0, stream, x,
                       N, fp, nl, sf, out);
    return hipGetLastError();
}

hipError_t launch_normalise(const float* x, int64_t N, float* out, float* work, int parts,
                            hipStream_t stream)
{
    float* partial = work;
    float* mean = work + 2 * parts;
    float* stdv = mean + 2;
    hipLaunchKernelGGL(moments_partial_kernel, dim3(parts), dim3(256), 0, stream, x, N, 0,
                       (const float*)nullptr, partial);
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, stream, partial, parts, N, 0,
                       mean);
    hipLaunchKernelGGL(moments_partial_kernel, dim3(parts), dim3(256), 0, stream, x, N, 1,
                       (const float*)mean, partial);
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, stream, partial, parts, N, 1,
                       stdv);
    hipLaunchKernelGGL(normalise_apply_kernel, dim3(grid_for((N + 3) / 4, 256)), dim3(256), 0, stream, x, N,
                       (const float*)mean, (const float*)stdv, out);
    return hipGetLastError();
}

hipError_t launch_prep_batch(const PrepSeriesDev* d_ps, int nseries, int64_t max_n,
                            int64_t max_lores, int max_width, int deredden, int normalise,
                            hipStream_t stream)
{
    if (nseries <= 0) return hipSuccess;
    if (deredden) {
        if (max_lores > 0) {
            hipLaunchKernelGGL(prep_scrunch_kernel, dim3(grid_for(max_lores, 64, 4096), nseries),
                               dim3(64), 0, stream, d_ps);
        }
        // the sorted path when its LDS fits in 48 KiB
        const size_t lds = rmed_lds_floats(256, max_width) * sizeof(float);
        const int use_lds = lds <= 48 * 1024;
        hipLaunchKernelGGL(prep_rmed_kernel,
                           dim3(grid_for(std::max(max_lores, (int64_t)1), 256, 4096), nseries),
                           dim3(256), use_lds ? lds : 0, stream, d_ps, use_lds);
        hipLaunchKernelGGL(prep_subtract_kernel, dim3(grid_for((max_n + 3) / 4, 256, 2048), nseries), dim3(256),
                           0, stream, d_ps);
    }
    if (normalise) {
        hipLaunchKernelGGL(prep_moments_partial_kernel, dim3(kNormaliseParts, nseries), dim3(256),
                           0, stream, d_ps, 0);
        hipLaunchKernelGGL(prep_moments_final_kernel, dim3(1, nseries), dim3(256), 0, stream, d_ps,
                           0);
        hipLaunchKernelGGL(prep_moments_partial_kernel, dim3(kNormaliseParts, nseries), dim3(256),
                           0, stream, d_ps, 1);
        hipLaunchKernelGGL(prep_moments_final_kernel, dim3(1, nseries), dim3(256), 0, stream, d_ps,
                           1);
        const int gx = grid_for((max_n + 3) / 4, 256);
        hipLaunchKernelGGL(prep_normalise_apply_kernel, dim3(gx, nseries), dim3(256), 0, stream,
                           d_ps, gx);
    }
    return hipGetLastError();
}

hipError_t launch_rollback(const float* x, int64_t size, int64_t shift, float* out,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(rollback_kernel, dim3(grid_for(size, 256)), dim3(256), 0, stream, x, size,
                       shift, out);
    return hipGetLastError();
}

hipError_t launch_fused_rollback_add(const float* x, const float* y, int64_t size, int64_t shift,
                                     float* z, hipStream_t stream)
{
    hipLaunchKernelGGL(fused_rollback_add_kernel, dim3(grid_for(size, 256)), dim3(256), 0, stream,
                       x, y, size, shift, z);
    return hipGetLastError();
}

hipError_t launch_circular_prefix_sum(const float* x, int64_t size, int64_t nsum, float* out,
                                      hipStream_t stream)
{
    hipLaunchKernelGGL(circular_prefix_sum_kernel, dim3(1), dim3(64), 0, stream, x, size, nsum,
                       out);
    return hipGetLastError();
}

hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace rt
