// Hand-written adjoint kernels of the periodogram path (normalise ->
// downsample -> fold -> FFA -> boxcar S/N), for gfx950.  The reference has no
// gradients; these are the transposes of the forward kernels' linear maps and
// the subgradient of the boxcar S/N at the forward's argmax window.
#include <hip/hip_runtime.h>

#include "adjoint.h"
#include "exact_f64.h"
#include "ffa_device.h"
#include "series_kernels.h"

namespace rt {

static inline int grid_of(int64_t n, int threads)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

// ---------------------------------------------------------------------------
// Boxcar S/N backward (snr.hpp:37-55): snr_w = ((h+b)*dmax_w - b*total)/std,
// dmax_w = cps[i*+w] - cps[i*] = sum of x over the window [i*+1, i*+w]
// (circular) for the first i* that attains the maximum.  d snr_w / d x[j] = ((h+b)*[j in window] - b)/std.
// One wavefront per row; rows with an all-zero upstream gradient get zeros.
// ---------------------------------------------------------------------------

// Rows [block * waves, (block + 1) * waves) of one transform, one row per
// wavefront; `lds` holds waves * bins floats plus the window tables.
__device__ __forceinline__ void snr_backward_block(
    const float* __restrict__ x, int rows, int bins, const float* __restrict__ gsnr,
    const int32_t* __restrict__ widths, const float* __restrict__ hb, int nw, float stdnoise,
    float* __restrict__ gx, int block, float* lds)
{
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = blockDim.x / kWave;
    const int r = block * nwaves + wave;
    float* cps = lds + (size_t)wave * bins;
    // per wave: window start and height increment of each width
    int* s_start = reinterpret_cast<int*>(lds + (size_t)nwaves * bins) + wave * kSnrBwdMaxW;
    float* s_up = reinterpret_cast<float*>(reinterpret_cast<int*>(lds + (size_t)nwaves * bins) +
                                           nwaves * kSnrBwdMaxW) + wave * kSnrBwdMaxW;
    const bool active = r < rows;
    if (active)
        for (int c = lane; c < bins; c += kWave) cps[c] = x[(size_t)r * bins + c];
    __syncthreads();
    if (active) snr_prefix_long(cps, cps, bins, lane);
    __syncthreads();
    if (!active) return;
    const float total = cps[bins - 1];
    float base = 0.0f;  // -sum_w g_w * b_w / std, added to every bin
    for (int iw = 0; iw < nw; ++iw) {
        const float gw = gsnr[(size_t)r * nw + iw];
        const int w = widths[iw];
        // first argmax of cps[i+w] - cps[i] (the reference keeps the first maximum)
        float best = -INFINITY;
        int bi = 0x7fffffff;
        if (gw != 0.0f) {
            for (int i = lane; i < bins; i += kWave) {
                const int j = i + w;
                const float v = (j < bins) ? cps[j] : cps[j - bins] + total;
                const float d = v - cps[i];
                if (d > best) {
                    best = d;
                    bi = i;
                }
            }
            for (int o = 32; o >= 1; o >>= 1) {
                const float ob = __shfl_xor(best, o, kWave);
                const int oi = __shfl_xor(bi, o, kWave);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
        }
        if (lane == 0) {
            // cps[i + w] - cps[i] sums x[i+1 .. i+w]: the window starts at i + 1
            s_start[iw] = (gw != 0.0f) ? (bi + 1 == bins ? 0 : bi + 1) : -1;
            s_up[iw] = gw * hb[2 * iw] / stdnoise;
        }
        base -= gw * hb[2 * iw + 1] / stdnoise;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float* g = gx + (size_t)r * bins;
    for (int c = lane; c < bins; c += kWave) {
        float v = base;
        for (int iw = 0; iw < nw; ++iw) {
            const int st = s_start[iw];
            if (st < 0) continue;
            int off = c - st;
            if (off < 0) off += bins;
            if (off < widths[iw]) v += s_up[iw];
        }
        g[c] = v;
    }
}

__global__ __launch_bounds__(256) void snr_backward_kernel(
    const float* __restrict__ x, int rows, int bins, const float* __restrict__ gsnr,
    const int32_t* __restrict__ widths, const float* __restrict__ hb, int nw, float stdnoise,
    float* __restrict__ gx)
{
    extern __shared__ float lds[];
    snr_backward_block(x, rows, bins, gsnr, widths, hb, nw, stdnoise, gx, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// FFA adjoint, one merge level (transforms.hpp:13-27 transposed): the
// forward adds head[h(s)] and roll(tail[t(s)], -(s - t(s))) into out[s]; the
// adjoint gathers, for every child row, the parent rows that read it.
// g_child has the state layout of depth d+1, g_parent of depth d.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void ffa_adjoint_level_range(const float* __restrict__ gp,
                                                        float* __restrict__ gc, int64_t rows,
                                                        int64_t cols, int depth, int64_t first,
                                                        int64_t stride)
{
    const int64_t total = rows * cols;
    for (int64_t e = first; e < total; e += stride) {
        const int64_t r = e / cols;
        const int64_t c = e - r * cols;
        // parent node at `depth` containing row r
        int64_t r0 = 0, m = rows;
        for (int d = 0; d < depth && m > 1; ++d) {
            const int64_t mh = m >> 1;
            if (r < r0 + mh) m = mh;
            else { r0 += mh; m -= mh; }
        }
        if (m <= 1) {
            gc[e] = gp[e];  // leaf above this level: passed through
            continue;
        }
        const int64_t mh = m >> 1;
        const int64_t mt = m - mh;
        const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
        const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
        const bool in_head = r < r0 + mh;
        const int64_t child = in_head ? r - r0 : r - r0 - mh;
        const float k = in_head ? kh : kt;
        // parent rows s with map(s) == child: a contiguous range around
        // child / k; scan a window and test exactly
        float acc = 0.0f;
        const int64_t guess = (k > 0.0f) ? (int64_t)((float)child / k) : 0;
        const int64_t s_lo = guess - 3 < 0 ? 0 : guess - 3;
        const int64_t s_hi = guess + 4 > m - 1 ? m - 1 : guess + 4;
        if (k == 0.0f) {
            // every parent row reads child 0 (two-row node)
            for (int64_t s = 0; s < m; ++s) {
                if (in_head) {
                    acc += gp[(r0 + s) * cols + c];
                } else {
                    int64_t cc = c - (s % cols);
                    if (cc < 0) cc += cols;
                    acc += gp[(r0 + s) * cols + cc];
                }
            }
        } else {
            for (int64_t s = s_lo; s <= s_hi; ++s) {
                const int mapped = merge_row(k, (int)s);
                if (mapped != child) continue;
                if (in_head) {
                    acc += gp[(r0 + s) * cols + c];
                } else {
                    // out[s][j] += tail[t][(j + sh) mod p]  =>  j = c - sh
                    const int64_t sh = (s - mapped) % cols;
                    int64_t cc = c - sh;
                    if (cc < 0) cc += cols;
                    acc += gp[(r0 + s) * cols + cc];
                }
            }
        }
        gc[e] = acc;
    }
}

__global__ void ffa_adjoint_level_kernel(const float* __restrict__ gp, float* __restrict__ gc,
                                         int64_t rows, int64_t cols, int depth)
{
    ffa_adjoint_level_range(gp, gc, rows, cols, depth,
                            (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                            (int64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Downsampling adjoint (downsample.hpp:58-78 transposed): every input sample
// collects the weights it had in the (at most three) outputs that read it.
// ---------------------------------------------------------------------------

// Index table of the downsampling by f: output k reads inputs [imin, imax]
// with edge weights wmin, wmax (the exact float64 index arithmetic of the
// reference, evaluated once per output).
__global__ void downsample_index_kernel(int64_t n, uint64_t mf, int ef, int64_t N,
                                        int4* __restrict__ idx)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const DsIndex ix = ds_index_fast(k, mf, ef, N);
        idx[k] = make_int4((int)ix.imin, (int)ix.imax, __float_as_int(ix.wmin),
                           __float_as_int(ix.wmax));
    }
}

// Every input sample i gathers the weights it had in the outputs that read
// it: those lie among k = floor(i/f) - 2 .. floor(i/f) + 3 (an output spans
// at least one input, its ends are floor(k*f) and floor(k*f + f) to within
// one sample, and i*(1/f) in float32 is within one of i/f).
__device__ __forceinline__ float downsample_adjoint_at(const float* __restrict__ gout, int64_t n,
                                                      const int4* __restrict__ idx, float inv_f,
                                                      int64_t i)
{
    int64_t k0 = (int64_t)((float)i * inv_f) - 2;
    if (k0 < 0) k0 = 0;
    float acc = 0.0f;
    for (int64_t k = k0; k < k0 + 6 && k < n; ++k) {
        const int4 ix = idx[k];
        if (i < ix.x || i > ix.y) continue;
        float w = 0.0f;
        if (i == ix.x) w += __int_as_float(ix.z);
        if (i == ix.y) w += __int_as_float(ix.w);
        if (i > ix.x && i < ix.y) w += 1.0f;
        acc += w * gout[k];
    }
    return acc;
}

__global__ void downsample_adjoint_kernel(const float* __restrict__ gout, int64_t n,
                                          const int4* __restrict__ idx, float inv_f, int64_t N,
                                          float* __restrict__ gin, int accumulate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float acc = downsample_adjoint_at(gout, n, idx, inv_f, i);
        gin[i] = accumulate ? gin[i] + acc : acc;
    }
}

// gin[i] (+)= gout[i] for i < n (fold of the raw series when f == 1)
__global__ void add_prefix_kernel(const float* __restrict__ gout, int64_t n, int64_t N,
                                  float* __restrict__ gin, int accumulate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = i < n ? gout[i] : 0.0f;
        gin[i] = accumulate ? gin[i] + v : v;
    }
}

// ---------------------------------------------------------------------------
// Normalisation adjoint: y = (x - m)/s with m, s the mean and std of x:
//   dx = (g - mean(g) - y * mean(g*y)) / s
// ---------------------------------------------------------------------------

__device__ __forceinline__ void normalise_adjoint_partial(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N, float* __restrict__ part,
    int block, int nblocks)
{
    __shared__ float sh[4][4];
    float a = 0.0f, al = 0.0f, b = 0.0f, bl = 0.0f;
    for (int64_t i = (int64_t)block * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)nblocks * blockDim.x) {
        ff_add(a, al, g[i], 0.0f, a, al);
        const float p = g[i] * y[i];
        ff_add(b, bl, p, __fmaf_rn(g[i], y[i], -p), b, bl);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const float oa = __shfl_down(a, d, kWave), oal = __shfl_down(al, d, kWave);
        const float ob = __shfl_down(b, d, kWave), obl = __shfl_down(bl, d, kWave);
        ff_add(a, al, oa, oal, a, al);
        ff_add(b, bl, ob, obl, b, bl);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    if (lane == 0) {
        sh[wave][0] = a; sh[wave][1] = al; sh[wave][2] = b; sh[wave][3] = bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); ++w) {
            ff_add(a, al, sh[w][0], sh[w][1], a, al);
            ff_add(b, bl, sh[w][2], sh[w][3], b, bl);
        }
        part[4 * block + 0] = a;
        part[4 * block + 1] = al;
        part[4 * block + 2] = b;
        part[4 * block + 3] = bl;
    }
}

__global__ __launch_bounds__(256) void normalise_adjoint_partial_kernel(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N, float* __restrict__ part)
{
    normalise_adjoint_partial(g, y, N, part, blockIdx.x, gridDim.x);
}

// The partial float-float sums are reduced by every block in parallel (lane
// p of the block takes partials p, p + blockDim, ...), then applied.
__device__ __forceinline__ void normalise_adjoint_apply(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N,
    const float* __restrict__ part, int nparts, float inv_std, float* __restrict__ gx, int block,
    int nblocks)
{
    __shared__ float sh[4][4];
    __shared__ float s_mg, s_mgy;
    float a = 0.0f, al = 0.0f, b = 0.0f, bl = 0.0f;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        ff_add(a, al, part[4 * p], part[4 * p + 1], a, al);
        ff_add(b, bl, part[4 * p + 2], part[4 * p + 3], b, bl);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const float oa = __shfl_down(a, d, kWave), oal = __shfl_down(al, d, kWave);
        const float ob = __shfl_down(b, d, kWave), obl = __shfl_down(bl, d, kWave);
        ff_add(a, al, oa, oal, a, al);
        ff_add(b, bl, ob, obl, b, bl);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    if (lane == 0) {
        sh[wave][0] = a; sh[wave][1] = al; sh[wave][2] = b; sh[wave][3] = bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); ++w) {
            ff_add(a, al, sh[w][0], sh[w][1], a, al);
            ff_add(b, bl, sh[w][2], sh[w][3], b, bl);
        }
        s_mg = (a + al) / (float)N;
        s_mgy = (b + bl) / (float)N;
    }
    __syncthreads();
    const float mg = s_mg, mgy = s_mgy;
    for (int64_t i = (int64_t)block * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)nblocks * blockDim.x)
        gx[i] = (g[i] - mg - y[i] * mgy) * inv_std;
}

__global__ __launch_bounds__(256) void normalise_adjoint_apply_kernel(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N,
    const float* __restrict__ part, int nparts, float inv_std, float* __restrict__ gx)
{
    normalise_adjoint_apply(g, y, N, part, nparts, inv_std, gx, blockIdx.x, gridDim.x);
}

// flags[t] = 1 if any of the rows_eval x nw S/N gradients of transform t
// (starting at row out_row) is non-zero; one block per transform.
__global__ __launch_bounds__(256) void transform_need_kernel(const float* __restrict__ gsnr,
                                                             const int2* __restrict__ rows,
                                                             int nw, int* __restrict__ flags)
{
    const int2 r = rows[blockIdx.x];
    const float* g = gsnr + (size_t)r.x * nw;
    const int64_t cnt = (int64_t)r.y * nw;
    int any = 0;
    for (int64_t e = threadIdx.x; e < cnt; e += blockDim.x) any |= g[e] != 0.0f;
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flags[blockIdx.x] = any;
}

// ---------------------------------------------------------------------------
// Batched backward over (series, transform) pairs and (series, step) groups
// of series that share one search plan.  Every kernel runs the one-series
// kernels' per-element arithmetic on the pair's or group's buffers, in the
// same order, so the results are bit-identical to one-series backwards.
// Blocks map to items through host-built tables: blk[b] = (item, local
// block, blocks of the item, depth).
// ---------------------------------------------------------------------------

// The buffer a pair's gradient ends in after `levels` transposed merge levels.
__device__ __forceinline__ const float* grad_pair_final(const GradPair& p, int levels)
{
    if (levels == 0) return p.A;
    return ((levels - 1) & 1) ? p.C : p.B;
}

__global__ void flags_batch_kernel(const float* __restrict__ gsnr, int64_t series_stride,
                                   const int2* __restrict__ rows, int ntr, int nw,
                                   int* __restrict__ flags)
{
    const int t = blockIdx.x, s = blockIdx.y;
    const int2 r = rows[t];
    const float* g = gsnr + (size_t)s * series_stride + (size_t)r.x * nw;
    const int64_t cnt = (int64_t)r.y * nw;
    int any = 0;
    for (int64_t e = threadIdx.x; e < cnt; e += blockDim.x) any |= g[e] != 0.0f;
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flags[(size_t)s * ntr + t] = any;
}

__global__ void ds_index_steps_kernel(const GradStep* __restrict__ steps,
                                      const int4* __restrict__ blk, int64_t N)
{
    const int4 b = blk[blockIdx.x];
    const GradStep st = steps[b.x];
    for (int64_t k = (int64_t)b.y * blockDim.x + threadIdx.x; k < st.n;
         k += (int64_t)b.z * blockDim.x) {
        const DsIndex ix = ds_index_fast(k, st.mf, st.ef, N);
        st.idx[k] = make_int4((int)ix.imin, (int)ix.imax, __float_as_int(ix.wmin),
                              __float_as_int(ix.wmax));
    }
}

__global__ void ds_values_groups_kernel(const GradGroup* __restrict__ groups,
                                        const GradStep* __restrict__ steps,
                                        const int4* __restrict__ blk)
{
    const int4 b = blk[blockIdx.x];
    const GradGroup g = groups[b.x];
    const GradStep st = steps[g.step];
    for (int64_t k = (int64_t)b.y * blockDim.x + threadIdx.x; k < st.n;
         k += (int64_t)b.z * blockDim.x) {
        const int4 e = st.idx[k];
        DsIndex ix;
        ix.imin = e.x;
        ix.imax = e.y;
        ix.wmin = __int_as_float(e.z);
        ix.wmax = __int_as_float(e.w);
        g.ds[k] = ds_accumulate(g.x, ix);
    }
}

// zeros in every pair's S/N-gradient rows (A), before the S/N backward
__global__ void fill_pairs_kernel(const GradPair* __restrict__ pairs, const int4* __restrict__ blk)
{
    const int4 b = blk[blockIdx.x];
    const GradPair p = pairs[b.x];
    const int64_t n = (int64_t)p.rows * p.bins;
    for (int64_t i = (int64_t)b.y * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)b.z * blockDim.x)
        p.A[i] = 0.0f;
}

__global__ __launch_bounds__(256) void snr_backward_pairs_kernel(
    const GradPair* __restrict__ pairs, const int4* __restrict__ blk,
    const int32_t* __restrict__ widths, int nw)
{
    extern __shared__ float lds[];
    const int4 b = blk[blockIdx.x];
    const GradPair p = pairs[b.x];
    snr_backward_block(p.T, p.rows_eval, p.bins, p.gsnr, widths, p.hb, nw, p.stdnoise, p.A, b.y,
                       lds);
}

__global__ void ffa_adjoint_level_pairs_kernel(const GradPair* __restrict__ pairs,
                                               const int4* __restrict__ blk)
{
    const int4 b = blk[blockIdx.x];
    const GradPair p = pairs[b.x];
    const int d = b.w;
    const float* src = d == 0 ? p.A : (((d - 1) & 1) ? p.C : p.B);
    float* dst = (d & 1) ? p.C : p.B;
    ffa_adjoint_level_range(src, dst, p.rows, p.bins, d,
                            (int64_t)b.y * blockDim.x + threadIdx.x, (int64_t)b.z * blockDim.x);
}

// gds of a (series, step) group: the fold transposes of its pairs' gradients
// added in pair order, from zero
__global__ void fold_groups_kernel(const GradGroup* __restrict__ groups,
                                   const GradPair* __restrict__ pairs,
                                   const GradStep* __restrict__ steps,
                                   const int4* __restrict__ blk)
{
    const int4 b = blk[blockIdx.x];
    const GradGroup g = groups[b.x];
    const int64_t n = steps[g.step].n;
    for (int64_t i = (int64_t)b.y * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)b.z * blockDim.x) {
        float v = 0.0f;
        for (int q = g.p0; q < g.p1; ++q) {
            const GradPair& p = pairs[q];
            const int64_t m = (int64_t)p.rows * p.bins;
            v = v + (i < m ? grad_pair_final(p, p.H)[i] : 0.0f);
        }
        g.gds[i] = v;
    }
}

// gradient of every series with respect to its (normalised) input: the
// downsampling transposes of its groups added in step order, from zero;
// series = blockIdx.y, its groups [g0[s], g0[s + 1])
__global__ void ds_adjoint_series_kernel(const GradGroup* __restrict__ groups,
                                         const GradStep* __restrict__ steps,
                                         const int* __restrict__ g0, int64_t N,
                                         float* __restrict__ gz)
{
    const int s = blockIdx.y;
    const int ga = g0[s], gb = g0[s + 1];
    float* out = gz + (size_t)s * N;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        float v = 0.0f;
        for (int q = ga; q < gb; ++q) {
            const GradGroup& g = groups[q];
            const GradStep& st = steps[g.step];
            const float c = st.idx ? downsample_adjoint_at(g.gds, st.n, st.idx, st.inv_f, i)
                                   : (i < st.n ? g.gds[i] : 0.0f);
            v = v + c;
        }
        out[i] = v;
    }
}

__global__ __launch_bounds__(256) void normalise_adjoint_partial_batch_kernel(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N, float* __restrict__ part,
    int nparts)
{
    const size_t s = blockIdx.y;
    normalise_adjoint_partial(g + s * N, y + s * N, N, part + s * 4 * nparts, blockIdx.x,
                              gridDim.x);
}

__global__ __launch_bounds__(256) void normalise_adjoint_apply_batch_kernel(
    const float* __restrict__ g, const float* __restrict__ y, int64_t N,
    const float* __restrict__ part, int nparts, const float* __restrict__ inv_std,
    float* __restrict__ gx)
{
    const size_t s = blockIdx.y;
    normalise_adjoint_apply(g + s * N, y + s * N, N, part + s * 4 * nparts, nparts, inv_std[s],
                            gx + s * N, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_flags_batch(const float* gsnr, int64_t series_stride, const int2* rows, int ntr,
                              int nw, int nseries, int* flags, hipStream_t s)
{
    if (ntr <= 0 || nseries <= 0) return hipSuccess;
    hipLaunchKernelGGL(flags_batch_kernel, dim3(ntr, nseries), dim3(256), 0, s, gsnr,
                       series_stride, rows, ntr, nw, flags);
    return hipGetLastError();
}

hipError_t launch_ds_index_steps(const GradStep* steps, const int4* blk, int nblk, int64_t N,
                                 hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(ds_index_steps_kernel, dim3(nblk), dim3(256), 0, s, steps, blk, N);
    return hipGetLastError();
}

hipError_t launch_ds_values_groups(const GradGroup* groups, const GradStep* steps,
                                   const int4* blk, int nblk, hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(ds_values_groups_kernel, dim3(nblk), dim3(256), 0, s, groups, steps, blk);
    return hipGetLastError();
}

hipError_t launch_fill_pairs(const GradPair* pairs, const int4* blk, int nblk, hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_pairs_kernel, dim3(nblk), dim3(256), 0, s, pairs, blk);
    return hipGetLastError();
}

hipError_t launch_snr_backward_pairs(const GradPair* pairs, const int4* blk, int nblk,
                                     int max_bins, const int32_t* widths, int nw, hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    if (nw > kSnrBwdMaxW) return hipErrorInvalidValue;
    const int waves = kGradSnrWaves;
    const size_t lds = (size_t)waves * max_bins * sizeof(float) + 2 * waves * kSnrBwdMaxW * 4;
    hipLaunchKernelGGL(snr_backward_pairs_kernel, dim3(nblk), dim3(64 * waves), lds, s, pairs, blk,
                       widths, nw);
    return hipGetLastError();
}

hipError_t launch_ffa_adjoint_level_pairs(const GradPair* pairs, const int4* blk, int nblk,
                                          hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(ffa_adjoint_level_pairs_kernel, dim3(nblk), dim3(256), 0, s, pairs, blk);
    return hipGetLastError();
}

hipError_t launch_fold_groups(const GradGroup* groups, const GradPair* pairs,
                              const GradStep* steps, const int4* blk, int nblk, hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(fold_groups_kernel, dim3(nblk), dim3(256), 0, s, groups, pairs, steps, blk);
    return hipGetLastError();
}

hipError_t launch_ds_adjoint_series(const GradGroup* groups, const GradStep* steps, const int* g0,
                                    int nseries, int64_t N, float* gz, hipStream_t s)
{
    if (nseries <= 0) return hipSuccess;
    hipLaunchKernelGGL(ds_adjoint_series_kernel, dim3(grid_of(N, 256), nseries), dim3(256), 0, s,
                       groups, steps, g0, N, gz);
    return hipGetLastError();
}

hipError_t launch_normalise_adjoint_batch(const float* g, const float* y, int64_t N, int nseries,
                                          const float* inv_std, float* part, int nparts,
                                          float* gx, hipStream_t s)
{
    if (nseries <= 0) return hipSuccess;
    hipLaunchKernelGGL(normalise_adjoint_partial_batch_kernel, dim3(nparts, nseries), dim3(256), 0,
                       s, g, y, N, part, nparts);
    hipLaunchKernelGGL(normalise_adjoint_apply_batch_kernel, dim3(grid_of(N, 256), nseries),
                       dim3(256), 0, s, g, y, N, (const float*)part, nparts, inv_std, gx);
    return hipGetLastError();
}


hipError_t launch_snr_backward(const float* x, int rows, int bins, const float* gsnr,
                               const int32_t* widths, const float* hb, int nw, float stdnoise,
                               float* gx, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    const int waves = 4;
    if (nw > kSnrBwdMaxW) return hipErrorInvalidValue;
    const size_t lds = (size_t)waves * bins * sizeof(float) + 2 * waves * kSnrBwdMaxW * 4;
    hipLaunchKernelGGL(snr_backward_kernel, dim3((rows + waves - 1) / waves), dim3(64 * waves),
                       lds, s, x, rows, bins, gsnr, widths, hb, nw, stdnoise, gx);
    return hipGetLastError();
}

hipError_t launch_ffa_adjoint_level(const float* gp, float* gc, int64_t rows, int64_t cols,
                                    int depth, hipStream_t s)
{
    hipLaunchKernelGGL(ffa_adjoint_level_kernel, dim3(grid_of(rows * cols, 256)), dim3(256), 0, s,
                       gp, gc, rows, cols, depth);
    return hipGetLastError();
}

hipError_t launch_downsample_adjoint(const float* gout, int64_t n, double f, int64_t N,
                                     float* gin, int accumulate, int4* idx, hipStream_t s)
{
    uint64_t mf;
    int ef;
    split_double(f, &mf, &ef);
    hipLaunchKernelGGL(downsample_index_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, n, mf, ef,
                       N, idx);
    hipLaunchKernelGGL(downsample_adjoint_kernel, dim3(grid_of(N, 256)), dim3(256), 0, s, gout, n,
                       (const int4*)idx, (float)(1.0 / f), N, gin, accumulate);
    return hipGetLastError();
}

hipError_t launch_add_prefix(const float* gout, int64_t n, int64_t N, float* gin, int accumulate,
                             hipStream_t s)
{
    hipLaunchKernelGGL(add_prefix_kernel, dim3(grid_of(N, 256)), dim3(256), 0, s, gout, n, N, gin,
                       accumulate);
    return hipGetLastError();
}

hipError_t launch_normalise_adjoint(const float* g, const float* y, int64_t N, float inv_std,
                                    float* part, int nparts, float* gx, hipStream_t s)
{
    hipLaunchKernelGGL(normalise_adjoint_partial_kernel, dim3(nparts), dim3(256), 0, s, g, y, N,
                       part);
    hipLaunchKernelGGL(normalise_adjoint_apply_kernel, dim3(grid_of(N, 256)), dim3(256), 0, s, g,
                       y, N, (const float*)part, nparts, inv_std, gx);
    return hipGetLastError();
}

hipError_t launch_transform_need(const float* gsnr, const int2* rows, int ntr, int nw, int* flags,
                                 hipStream_t s)
{
    if (ntr <= 0) return hipSuccess;
    hipLaunchKernelGGL(transform_need_kernel, dim3(ntr), dim3(256), 0, s, gsnr, rows, nw, flags);
    return hipGetLastError();
}

}  // namespace rt
