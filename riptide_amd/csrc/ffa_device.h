// Device helpers shared by the FFA / S/N kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include "tasks.h"

namespace rt {

// Register-resident profile positions per lane: rows up to 64*kSnrMaxK bins
// for the standalone S/N kernel, 64*kSnrFfaK inside the FFA kernel.
constexpr int kSnrMaxK = kSnrRegsMaxK;
constexpr int kSnrFfaK = 9;

// Child row of the merge map, trunc(k*s + 0.5f) (transforms.hpp:17-18).  The
// reference build contracts k*s + 0.5f into one fused multiply-add; so do we.
__device__ __forceinline__ int merge_row(float k, int s)
{
    return (int)__fmaf_rn(k, (float)s, 0.5f);
}

// Global-memory views of plain pointers: loads and stores through them are
// global_* instructions rather than flat ones (flat accesses also count
// against the LDS wait counter).
typedef __attribute__((address_space(1))) const float gcfloat;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const int32_t gcint;
__device__ __forceinline__ gcfloat* as_global(const float* p) { return (gcfloat*)p; }
__device__ __forceinline__ gfloat* as_global(float* p) { return (gfloat*)p; }
__device__ __forceinline__ gcint* as_global(const int32_t* p) { return (gcint*)p; }

// Error-free sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (ah, al) + (bh, bl) as a normalised float-float pair.
__device__ __forceinline__ void ff_add(float ah, float al, float bh, float bl, float& rh,
                                       float& rl)
{
    float s, e;
    two_sum(ah, bh, s, e);
    e += al + bl;
    rh = s + e;
    rl = e - (rh - s);
}

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP moves (no
// LDS traffic): shifts by 1, 2, 4, 8 lanes within each 16-lane row, then the
// row totals broadcast into the rows above.  A lane whose DPP source lies
// outside the pattern (or in a row the mask leaves out) adds 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_add(int v)
{
    return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, false);
}

__device__ __forceinline__ int wave_inclusive_sum(int v)
{
    v = dpp_add<0x111, 0xF>(v);  // row_shr:1
    v = dpp_add<0x112, 0xF>(v);  // row_shr:2
    v = dpp_add<0x114, 0xF>(v);  // row_shr:4
    v = dpp_add<0x118, 0xF>(v);  // row_shr:8
    v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v = fmaxf(v, __shfl_xor(v, d, kWave));
    return v;
}

// Circular prefix sum of one profile, in place (kernels.hpp:73-101): the
// reference accumulates in float64 and rounds each partial sum to float32;
// here a float-float (two-sum) scan across the wavefront gives the same
// float32 partial sums.  Lane l owns positions [l*k, l*k + k), k = ceil(bins/64).
// `own` receives the lane's partial sums.  Call with the whole wavefront; the
// caller must barrier before other lanes read the row.
template <int K>
__device__ __forceinline__ void snr_prefix_inplace(float* row, int bins, int lane, float (&own)[K])
{
    const int k = (bins + kWave - 1) / kWave;
    const int base = lane * k;
    float ph[K], pl[K];
    float sh = 0.0f, sl = 0.0f;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (q < k) {
            const int i = base + q;
            const float v = (i < bins) ? row[i] : 0.0f;
            float s, e;
            two_sum(sh, v, s, e);
            sh = s;
            sl += e;
            ph[q] = sh;
            pl[q] = sl;
        }
    }
    // inclusive scan of the lane totals
    float ih = sh, il = sl;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float oh = __shfl_up(ih, d, kWave);
        const float ol = __shfl_up(il, d, kWave);
        if (lane >= d)
            ff_add(oh, ol, ih, il, ih, il);
    }
    // exclusive prefix of this lane
    float eh = __shfl_up(ih, 1, kWave);
    float el = __shfl_up(il, 1, kWave);
    if (lane == 0) {
        eh = 0.0f;
        el = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (q < k) {
            float rh, rl;
            ff_add(eh, el, ph[q], pl[q], rh, rl);
            own[q] = rh;
            const int i = base + q;
            if (i < bins) row[i] = rh;
        }
    }
}

// Per-width boxcar S/N of a profile whose circular prefix sums are in `cps`
// (snr.hpp:44-54): dmax = max_i cps[i+w] - cps[i] over the circular
// extension cps[bins + j] = cps[j] + total (float32, kernels.hpp:100), then
// ((h+b)*dmax - b*total)/stdnoise with the reference build's contraction.
template <int K>
__device__ __forceinline__ void snr_widths(const float* cps, int bins, int lane,
                                           const float (&own)[K],
                                           const int32_t* widths, const float* hb, int nw,
                                           float stdnoise, float* out)
{
    const int k = (bins + kWave - 1) / kWave;
    const int base = lane * k;
    const float total = cps[bins - 1];
    float res = 0.0f;
    for (int iw = 0; iw < nw; ++iw) {
        const int w = widths[iw];
        float dm = -INFINITY;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (q < k) {
                const int i = base + q;
                if (i < bins) {
                    const int j = i + w;
                    const float v = (j < bins) ? cps[j] : cps[j - bins] + total;
                    dm = fmaxf(dm, v - own[q]);
                }
            }
        }
        dm = wave_max(dm);
        const float hpb = hb[2 * iw];
        const float bl = hb[2 * iw + 1];
        const float snr = __fmaf_rn(dm, hpb, -(bl * total)) / stdnoise;
        if ((iw & (kWave - 1)) == lane)
            res = snr;
        if ((iw & (kWave - 1)) == kWave - 1 || iw == nw - 1) {
            const int first = iw & ~(kWave - 1);
            if (first + lane <= iw)
                out[first + lane] = res;
        }
    }
}

// Long-row variant of the prefix sum: chunks of 64*4 positions, the running
// float-float carry broadcast between chunks, results written to `cps`.
__device__ __forceinline__ void snr_prefix_long(const float* x, float* cps, int bins, int lane)
{
    constexpr int K = 4;
    float ch = 0.0f, cl = 0.0f;  // carry: sum of all previous chunks
    for (int c0 = 0; c0 < bins; c0 += kWave * K) {
        const int base = c0 + lane * K;
        float ph[K], pl[K];
        float sh = 0.0f, sl = 0.0f;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int i = base + q;
            const float v = (i < bins) ? x[i] : 0.0f;
            float s, e;
            two_sum(sh, v, s, e);
            sh = s;
            sl += e;
            ph[q] = sh;
            pl[q] = sl;
        }
        float ih = sh, il = sl;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const float oh = __shfl_up(ih, d, kWave);
            const float ol = __shfl_up(il, d, kWave);
            if (lane >= d)
                ff_add(oh, ol, ih, il, ih, il);
        }
        float eh = __shfl_up(ih, 1, kWave);
        float el = __shfl_up(il, 1, kWave);
        if (lane == 0) {
            eh = 0.0f;
            el = 0.0f;
        }
        ff_add(ch, cl, eh, el, eh, el);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int i = base + q;
            float rh, rl;
            ff_add(eh, el, ph[q], pl[q], rh, rl);
            if (i < bins) cps[i] = rh;
        }
        const float th = __shfl(ih, kWave - 1, kWave);
        const float tl = __shfl(il, kWave - 1, kWave);
        ff_add(ch, cl, th, tl, ch, cl);
    }
}

__device__ __forceinline__ void snr_widths_long(const float* cps, int bins, int lane,
                                                const int32_t* widths, const float* hb, int nw,
                                                float stdnoise, float* out)
{
    const float total = cps[bins - 1];
    for (int iw = 0; iw < nw; ++iw) {
        const int w = widths[iw];
        float dm = -INFINITY;
        for (int i = lane; i < bins; i += kWave) {
            const int j = i + w;
            const float v = (j < bins) ? cps[j] : cps[j - bins] + total;
            dm = fmaxf(dm, v - cps[i]);
        }
        dm = wave_max(dm);
        if (lane == 0)
            out[iw] = __fmaf_rn(dm, hb[2 * iw], -(hb[2 * iw + 1] * total)) / stdnoise;
    }
}

}  // namespace rt
