// Boxcar S/N and global-memory FFA kernels for gfx950, and the dispatch of
// the FFA task kernels (ffa_task.h) to their compiled variants.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ffa_device.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "tasks.h"

namespace rt {


// Plain S/N of rows x cols profiles (python_bindings.cpp:121-160): one
// wavefront per row, the row staged in LDS.
__global__ __launch_bounds__(256) void snr_rows_kernel(const float* __restrict__ x, int rows,
                                                       int cols, const int32_t* widths,
                                                       const float* hb, int nw,
                                                       float stdnoise, float* __restrict__ out)
{
    extern __shared__ float lds[];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = blockDim.x / kWave;
    const int r = blockIdx.x * nwaves + wave;
    float* row = lds + (size_t)wave * cols;
    if (r < rows) {
        const float* src = x + (size_t)r * cols;
        for (int c = lane; c < cols; c += kWave)
            row[c] = src[c];
    }
    float own[kSnrMaxK];
    __syncthreads();
    if (r < rows)
        snr_prefix_inplace<kSnrMaxK>(row, cols, lane, own);
    __syncthreads();
    if (r < rows)
        snr_widths<kSnrMaxK>(row, cols, lane, own, widths, hb, nw, stdnoise, out + (size_t)r * nw);
}

// Same for rows longer than the register path supports: the prefix sums go
// to a global scratch row, one workgroup (one wavefront) per row.
__global__ __launch_bounds__(64) void snr_rows_long_kernel(const float* __restrict__ x, int rows,
                                                           int cols, const int32_t* widths,
                                                           const float* hb, int nw,
                                                           float stdnoise, float* scratch,
                                                           float* __restrict__ out)
{
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    if (r >= rows) return;
    float* cps = scratch + (size_t)r * cols;
    snr_prefix_long(x + (size_t)r * cols, cps, cols, lane);
    __syncthreads();
    snr_widths_long(cps, cols, lane, widths, hb, nw, stdnoise, out + (size_t)r * nw);
}

// One FFA merge level straight from HBM for blocks too wide for LDS: every
// output element adds the two operands the reference adds.  Launched once
// per tree depth, deepest first.
__global__ void ffa_level_global_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                        int64_t rows, int64_t cols, int depth)
{
    const int64_t total = rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = e / cols;
        const int64_t c = e - p * cols;
        // node at `depth` containing row p
        int64_t r0 = 0, m = rows;
        for (int d = 0; d < depth && m > 1; ++d) {
            const int64_t mh = m >> 1;
            if (p < r0 + mh) m = mh;
            else { r0 += mh; m -= mh; }
        }
        if (m <= 1) {
            dst[e] = src[e];
            continue;
        }
        const int64_t mh = m >> 1;
        const int64_t mt = m - mh;
        const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
        const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
        const int64_t s = p - r0;
        const int64_t h = merge_row(kh, (int)s);
        const int64_t t = merge_row(kt, (int)s);
        int64_t cc = c + (s - t) % cols;
        if (cc >= cols) cc -= cols;
        dst[e] = src[(r0 + h) * cols + c] + src[(r0 + mh + t) * cols + cc];
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Dispatch of the FFA task kernels to their variants
// ---------------------------------------------------------------------------

hipError_t launch_ffa_tasks(const FfaTask* d_tasks, int ntasks, int* d_err, hipStream_t stream,
                            int shape, int variant)
{
    if (ntasks <= 0) return hipSuccess;
    switch (variant) {
#define RT_V(V) \
    case V: return launch_ffa_tasks_variant<V>(d_tasks, ntasks, d_err, stream, shape);
    RT_V(0) RT_V(1) RT_V(2) RT_V(3) RT_V(4) RT_V(5) RT_V(6) RT_V(7) RT_V(8) RT_V(9)
#undef RT_V
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ffa_task_list(const FfaTask* d_tasks, const FfaTask* h_tasks, int ntasks,
                                int* d_err, hipStream_t stream, int shape)
{
    for (int i = 0; i < ntasks;) {
        const int v = ffa_variant(h_tasks[i].bins);
        int j = i + 1;
        while (j < ntasks && ffa_variant(h_tasks[j].bins) == v) ++j;
        const hipError_t e = launch_ffa_tasks(d_tasks + i, j - i, d_err, stream, shape, v);
        if (e != hipSuccess) return e;
        i = j;
    }
    return hipSuccess;
}

void sort_by_variant(std::vector<FfaTask>& tasks)
{
    std::stable_sort(tasks.begin(), tasks.end(), [](const FfaTask& a, const FfaTask& b) {
        return ffa_variant(a.bins) < ffa_variant(b.bins);
    });
}

hipError_t launch_ffa_batch(const FfaTransform* d_trs, int ntr, int* d_queue, float* d_scratch,
                            int64_t slot_floats, int slots, int* d_err, hipStream_t stream,
                            int shape, int variant)
{
    if (ntr <= 0) return hipSuccess;
    switch (variant) {
#define RT_V(V)                                                                              \
    case V:                                                                                  \
        return launch_ffa_batch_variant<V>(d_trs, ntr, d_queue, d_scratch, slot_floats, slots, \
                                           d_err, stream, shape);
    RT_V(0) RT_V(1) RT_V(2) RT_V(3) RT_V(4) RT_V(5) RT_V(6) RT_V(7) RT_V(8) RT_V(9)
#undef RT_V
    default: return hipErrorInvalidValue;
    }
}
static_assert(kFfaVariants == 10, "dispatch covers variants 0..9");

hipError_t launch_snr_rows(const float* x, int64_t rows, int cols, const int32_t* widths,
                           const float* hb, int nw, float stdnoise, float* out,
                           float* long_scratch, hipStream_t stream)
{
    if (rows <= 0) return hipSuccess;
    if (cols <= kWave * kSnrMaxK) {
        const int waves = 4;
        hipLaunchKernelGGL(snr_rows_kernel, dim3((unsigned)((rows + waves - 1) / waves)),
                           dim3(waves * kWave), (size_t)waves * cols * sizeof(float), stream, x,
                           (int)rows, cols, widths, hb, nw, stdnoise, out);
    } else {
        hipLaunchKernelGGL(snr_rows_long_kernel, dim3((unsigned)rows), dim3(kWave), 0, stream, x,
                           (int)rows, cols, widths, hb, nw, stdnoise, long_scratch, out);
    }
    return hipGetLastError();
}

hipError_t launch_ffa_level_global(const float* src, float* dst, int64_t rows, int64_t cols,
                                   int depth, hipStream_t stream)
{
    const int64_t total = rows * cols;
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(ffa_level_global_kernel, dim3((unsigned)g), dim3(256), 0, stream, src, dst,
                       rows, cols, depth);
    return hipGetLastError();
}

}  // namespace rt
