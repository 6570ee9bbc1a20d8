// FIR whitening filters of a batch of series, their adjoints and the Adam
// update of the filter taps: the differentiable front end of a batched search
// whose per-series filters are optimised for S/N.  The reference has no
// counterpart (it dereddens with a running median only, time_series.py:95-124).
//
// Filter of series s with T taps (centre c = (T - 1) / 2), zero outside the
// series:  y[s][n] = sum_k h[s][k] * x[s][n + k - c].
// Adjoints:  dL/dx[s][m] = sum_k h[s][k] * g[s][m - k + c]   (the same filter
//            with the taps reversed, T odd), and
//            dL/dh[s][k] = sum_n g[s][n] * x[s][n + k - c].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/riptide_amd.h"

namespace rt {

void set_error(const std::string& msg);

namespace {

int ffail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

#define FIR_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ffail(RT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kFirThreads = 256;
constexpr int kFirOut = 8;                          // outputs per thread
constexpr int kFirTile = kFirThreads * kFirOut;     // outputs per workgroup
constexpr int kFirMaxTaps = 1023;
constexpr int kGradTile = 4096;                     // samples per tap-gradient workgroup

// One workgroup per (tile of kFirTile outputs, series).  The tile's inputs
// with a halo of T - 1 samples are staged in LDS; thread t computes outputs
// t + 256 j, so the lanes of a wavefront read consecutive LDS words; the taps
// are uniform and come through the scalar cache.  `flip` applies the taps in
// reverse order (the input adjoint).
__global__ __launch_bounds__(kFirThreads) void fir_forward_kernel(const float* __restrict__ x,
                                                                  int64_t n,
                                                                  const float* __restrict__ h,
                                                                  int ntaps, int flip,
                                                                  float* __restrict__ y)
{
    extern __shared__ float xs[];  // kFirTile + ntaps - 1
    const int64_t s = blockIdx.y;
    const int64_t n0 = (int64_t)blockIdx.x * kFirTile;
    const int c = (ntaps - 1) / 2;
    const float* xr = x + s * n;
    const float* hr = h + s * ntaps;
    const int span = kFirTile + ntaps - 1;
    for (int i = threadIdx.x; i < span; i += kFirThreads) {
        const int64_t g = n0 + i - c;
        xs[i] = (g >= 0 && g < n) ? xr[g] : 0.0f;
    }
    __syncthreads();
    float acc[kFirOut];
#pragma unroll
    for (int j = 0; j < kFirOut; ++j) acc[j] = 0.0f;
    const int t = threadIdx.x;
    for (int k = 0; k < ntaps; ++k) {
        const float hk = hr[flip ? ntaps - 1 - k : k];
#pragma unroll
        for (int j = 0; j < kFirOut; ++j) acc[j] = __fmaf_rn(hk, xs[t + j * kFirThreads + k], acc[j]);
    }
    float* yr = y + s * n;
#pragma unroll
    for (int j = 0; j < kFirOut; ++j) {
        const int64_t o = n0 + t + j * kFirThreads;
        if (o < n) yr[o] = acc[j];
    }
}

// Partial tap gradients of one tile of kGradTile samples of series s:
// part[(s * tiles + tile) * T + k] = sum over the tile of g[n] x[n + k - c].
// Thread k owns tap k (lanes read consecutive x words, g is broadcast).
__global__ void fir_grad_taps_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                     int64_t n, int ntaps, int tiles, float* __restrict__ part)
{
    extern __shared__ float lds[];
    float* gs = lds;               // kGradTile
    float* xs = lds + kGradTile;   // kGradTile + ntaps - 1
    const int64_t s = blockIdx.y;
    const int tile = blockIdx.x;
    const int64_t n0 = (int64_t)tile * kGradTile;
    const int c = (ntaps - 1) / 2;
    const float* xr = x + s * n;
    const float* gr = g + s * n;
    const int span = kGradTile + ntaps - 1;
    for (int i = threadIdx.x; i < kGradTile; i += blockDim.x) {
        const int64_t q = n0 + i;
        gs[i] = q < n ? gr[q] : 0.0f;
    }
    for (int i = threadIdx.x; i < span; i += blockDim.x) {
        const int64_t q = n0 + i - c;
        xs[i] = (q >= 0 && q < n) ? xr[q] : 0.0f;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ntaps; k += blockDim.x) {
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        for (int i = 0; i < kGradTile; i += 4) {
            a0 = __fmaf_rn(gs[i], xs[i + k], a0);
            a1 = __fmaf_rn(gs[i + 1], xs[i + 1 + k], a1);
            a2 = __fmaf_rn(gs[i + 2], xs[i + 2 + k], a2);
            a3 = __fmaf_rn(gs[i + 3], xs[i + 3 + k], a3);
        }
        part[((size_t)s * tiles + tile) * ntaps + k] = (a0 + a1) + (a2 + a3);
    }
}

// gh[s][k] = sum over tiles of the partial sums, in tile order.
__global__ void fir_grad_reduce_kernel(const float* __restrict__ part, int64_t nseries,
                                       int ntaps, int tiles, float* __restrict__ gh)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nseries * ntaps) return;
    const int64_t s = e / ntaps;
    const int k = (int)(e - s * ntaps);
    const float* p = part + (size_t)s * tiles * ntaps + k;
    float acc = 0.0f;
    for (int t = 0; t < tiles; ++t) acc += p[(size_t)t * ntaps];
    gh[e] = acc;
}

// Adam (Kingma & Ba 2015, algorithm 1) on `count` parameters, step >= 1:
// m <- b1 m + (1 - b1) g, v <- b2 v + (1 - b2) g^2,
// p <- p - lr (m / (1 - b1^step)) / (sqrt(v / (1 - b2^step)) + eps).
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t count, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int fir_check(int64_t nseries, int64_t size, int64_t ntaps)
{
    if (nseries < 1 || size < 1) return ffail(RT_EINVAL, "FIR needs at least one non-empty series");
    if (ntaps < 1 || ntaps > kFirMaxTaps || ntaps % 2 == 0)
        return ffail(RT_EINVAL, "FIR tap count must be odd and at most 1023");
    if (nseries > 65535) return ffail(RT_EINVAL, "at most 65535 series per FIR call");
    return RT_OK;
}

int64_t grad_tiles(int64_t size) { return (size + kGradTile - 1) / kGradTile; }

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

int rt_fir_forward(const float* d_x, int64_t nseries, int64_t size, const float* d_taps,
                   int64_t ntaps, int reverse, float* d_y, void* stream)
{
    int st = fir_check(nseries, size, ntaps);
    if (st) return st;
    const int64_t tiles = (size + kFirTile - 1) / kFirTile;
    if (tiles > 0x7fffffff) return ffail(RT_EINVAL, "series too long for the FIR kernel");
    const size_t lds = (size_t)(kFirTile + ntaps - 1) * sizeof(float);
    hipLaunchKernelGGL(fir_forward_kernel, dim3((unsigned)tiles, (unsigned)nseries),
                       dim3(kFirThreads), lds, (hipStream_t)stream, d_x, size, d_taps, (int)ntaps,
                       reverse ? 1 : 0, d_y);
    FIR_HIP(hipGetLastError());
    return RT_OK;
}

int rt_fir_workspace_floats(int64_t nseries, int64_t size, int64_t ntaps, int64_t* floats)
{
    int st = fir_check(nseries, size, ntaps);
    if (st) return st;
    *floats = nseries * grad_tiles(size) * ntaps;
    return RT_OK;
}

int rt_fir_grad_taps(const float* d_x, const float* d_g, int64_t nseries, int64_t size,
                     int64_t ntaps, float* d_work, float* d_gtaps, void* stream)
{
    int st = fir_check(nseries, size, ntaps);
    if (st) return st;
    const int64_t tiles = grad_tiles(size);
    if (tiles > 0x7fffffff) return ffail(RT_EINVAL, "series too long for the FIR kernel");
    const int threads = (int)std::min<int64_t>(1024, ((ntaps + 63) / 64) * 64);
    const size_t lds = (size_t)(2 * kGradTile + ntaps - 1) * sizeof(float);
    hipLaunchKernelGGL(fir_grad_taps_kernel, dim3((unsigned)tiles, (unsigned)nseries),
                       dim3(threads), lds, (hipStream_t)stream, d_x, d_g, size, (int)ntaps,
                       (int)tiles, d_work);
    FIR_HIP(hipGetLastError());
    const int64_t total = nseries * ntaps;
    hipLaunchKernelGGL(fir_grad_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, d_work, nseries, (int)ntaps, (int)tiles, d_gtaps);
    FIR_HIP(hipGetLastError());
    return RT_OK;
}

int rt_adam_update(float* d_params, const float* d_grad, float* d_m, float* d_v, int64_t count,
                   double lr, double beta1, double beta2, double eps, int64_t step, void* stream)
{
    if (count < 1) return RT_OK;
    if (step < 1) return ffail(RT_EINVAL, "Adam step counts from 1");
    const double c1 = 1.0 - std::pow(beta1, (double)step);
    const double c2 = 1.0 - std::pow(beta2, (double)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, d_params, d_grad, d_m, d_v, count, (float)lr,
                       (float)beta1, (float)beta2, (float)eps, (float)c1, (float)c2);
    FIR_HIP(hipGetLastError());
    return RT_OK;
}

}  // extern "C"
