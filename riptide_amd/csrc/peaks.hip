// Peak detection on the periodogram S/N array without copying it to the host
// (peak_detection.py:38-240 of the reference, run on the host there):
//
//  1. segment statistics: the S/N of every width, cut into m consecutive
//     segments of p periods, needs the 25/50/75th percentiles of each
//     segment.  rt_segment_select returns, per (segment, width), the sorted
//     values at the ranks the percentile interpolation reads (exact order
//     statistics by a 4-pass radix select on order-preserving keys); the
//     interpolation itself and the polynomial fit stay on the host.
//  2. candidates: rt_peak_candidates compares every S/N with smin and with
//     the fitted threshold polynomial evaluated in float32; values farther
//     than a rounding bound from the threshold are decided here, the few
//     within it are returned separately for an exact float64 decision on the
//     host.  Clustering of the (few) selected periods stays on the host.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/riptide_amd.h"

namespace rt {

void set_error(const std::string& msg);

namespace {

int pfail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

#define PK_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return pfail(RT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kSelThreads = 256;
constexpr int kMaxRanks = 8;
constexpr int kMaxPolyDeg = 8;

// Target ranks of one selection, passed by value.
struct SelRanks {
    int32_t r[kMaxRanks];
};

// Order-preserving map of a float to an unsigned int (ascending keys =
// ascending floats), and back.
__device__ __forceinline__ uint32_t float_key(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// One workgroup per (segment j, width iw): the values of rank ranks[t]
// (0-based, ascending) among s[j*p + i][iw], i < p, for t < nranks.  Each of
// the 4 passes fixes 8 more bits of every target's key from a 256-bin
// histogram of the values that share the bits fixed so far.
__global__ __launch_bounds__(kSelThreads) void segment_select_kernel(
    const float* __restrict__ snrs, int64_t nw, int64_t p, SelRanks ranks, int nranks,
    float* __restrict__ out)
{
    __shared__ uint32_t hist[kMaxRanks][256];
    __shared__ uint32_t s_prefix[kMaxRanks];
    __shared__ uint32_t s_rank[kMaxRanks];
    const int64_t j = blockIdx.x;
    const int64_t iw = blockIdx.y;
    const float* col = snrs + (size_t)j * p * nw + iw;
    if (threadIdx.x < nranks) {
        s_prefix[threadIdx.x] = 0;
        s_rank[threadIdx.x] = (uint32_t)ranks.r[threadIdx.x];
    }
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        const uint32_t hi = pass == 0 ? 0u : (0xFFFFFFFFu << (32 - 8 * pass));
        for (int e = threadIdx.x; e < kMaxRanks * 256; e += kSelThreads) (&hist[0][0])[e] = 0;
        __syncthreads();
        uint32_t pre[kMaxRanks];
#pragma unroll
        for (int t = 0; t < kMaxRanks; ++t) pre[t] = t < nranks ? (s_prefix[t] & hi) : 0;
        for (int64_t i = threadIdx.x; i < p; i += kSelThreads) {
            const uint32_t key = float_key(col[(size_t)i * nw]);
            const uint32_t d = (key >> shift) & 255u;
#pragma unroll
            for (int t = 0; t < kMaxRanks; ++t)
                if (t < nranks && (key & hi) == pre[t]) atomicAdd(&hist[t][d], 1u);
        }
        __syncthreads();
        if (threadIdx.x < nranks) {
            const int t = threadIdx.x;
            uint32_t r = s_rank[t], cum = 0;
            int d = 0;
            for (; d < 255; ++d) {
                if (cum + hist[t][d] > r) break;
                cum += hist[t][d];
            }
            s_rank[t] = r - cum;
            s_prefix[t] = (s_prefix[t] & hi) | ((uint32_t)d << shift);
        }
        __syncthreads();
    }
    if (threadIdx.x < nranks)
        out[((size_t)j * nw + iw) * nranks + threadIdx.x] = key_float(s_prefix[threadIdx.x]);
}

// Every (period ip, width iw): s = snrs[ip][iw].  s must exceed smin_lo (the
// largest float32 <= smin, so that s > smin_lo <=> s > smin for a float32 s)
// and the threshold thr = poly_iw(log f[ip]) (coefficients highest degree
// first).  thr is evaluated in float32 with a bound `tol` on its distance
// from the float64 value; s > thr + tol is a sure pass (sel), |s - thr| <=
// tol undecided (unc), both appended per width (order not kept).
__global__ void peak_candidates_kernel(const float* __restrict__ snrs, int64_t L, int64_t nw,
                                       const float* __restrict__ logf,
                                       const float* __restrict__ coef, int ncoef, float smin_lo,
                                       int32_t* __restrict__ counts, int32_t* __restrict__ sel,
                                       int32_t* __restrict__ unc, int64_t cap)
{
    const int64_t total = L * nw;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const float s = snrs[e];
        if (!(s > smin_lo)) continue;
        const int64_t ip = e / nw;
        const int iw = (int)(e - ip * nw);
        const float x = logf[ip];
        const float ax = fabsf(x);
        const float* c = coef + (size_t)iw * ncoef;
        float thr = 0.0f, mag = 0.0f;
        for (int k = 0; k < ncoef; ++k) {
            thr = thr * x + c[k];
            mag = mag * ax + fabsf(c[k]);
        }
        const float tol = 4e-6f * mag + 1e-6f;
        if (s > thr + tol) {
            const int32_t q = atomicAdd(&counts[2 * iw], 1);
            if (q < cap) sel[(size_t)iw * cap + q] = (int32_t)ip;
        } else if (s >= thr - tol) {
            const int32_t q = atomicAdd(&counts[2 * iw + 1], 1);
            if (q < cap) unc[(size_t)iw * cap + q] = (int32_t)ip;
        }
    }
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

int rt_segment_select(const float* d_snrs, int64_t nw, int64_t nseg, int64_t seglen,
                      const int32_t* h_ranks, int64_t nranks, float* d_out, void* stream)
{
    if (nw < 1 || nseg < 1 || seglen < 1) return pfail(RT_EINVAL, "empty segment selection");
    if (nranks < 1 || nranks > kMaxRanks) return pfail(RT_EINVAL, "1 to 8 ranks per segment");
    if (nseg > 0x7fffffff || nw > 65535) return pfail(RT_EINVAL, "too many segments or widths");
    for (int64_t t = 0; t < nranks; ++t)
        if (h_ranks[t] < 0 || h_ranks[t] >= seglen) return pfail(RT_EINVAL, "rank outside the segment");
    SelRanks r = {};
    for (int64_t t = 0; t < nranks; ++t) r.r[t] = h_ranks[t];
    hipLaunchKernelGGL(segment_select_kernel, dim3((unsigned)nseg, (unsigned)nw), dim3(kSelThreads),
                       0, (hipStream_t)stream, d_snrs, nw, seglen, r, (int)nranks, d_out);
    PK_HIP(hipGetLastError());
    return RT_OK;
}

int rt_peak_candidates(const float* d_snrs, int64_t length, int64_t nw, const float* d_logf,
                       const float* d_coef, int64_t ncoef, float smin_lo, int32_t* d_counts,
                       int32_t* d_sel, int32_t* d_unc, int64_t cap, void* stream)
{
    if (length < 1 || nw < 1) return RT_OK;
    if (ncoef < 1 || ncoef > kMaxPolyDeg + 1)
        return pfail(RT_EINVAL, "threshold polynomial degree must be 0 to 8");
    if (cap < 1) return pfail(RT_EINVAL, "candidate capacity must be positive");
    int64_t blocks = (length * nw + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(peak_candidates_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, d_snrs, length, nw, d_logf, d_coef, (int)ncoef, smin_lo,
                       d_counts, d_sel, d_unc, cap);
    PK_HIP(hipGetLastError());
    return RT_OK;
}

}  // extern "C"
