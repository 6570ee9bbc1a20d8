// Device descriptors and host launchers of the per-series kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

constexpr int kDsThreads = 512;
constexpr int kDsMaxSteps = 256;
// Input samples staged per workgroup by the packed downsampling kernel, and
// the largest factor it handles (its halo, staged next to the chunk with
// their two prefix arrays, must fit in LDS); larger factors use the direct
// kernel.
constexpr int kDsChunk = 4096;
constexpr double kDsPackedMaxF = 4000.0;
// Staged samples per thread loaded together (a chunk plus a halo of up to
// 1024 samples; a longer halo takes a loop for the rest).
constexpr int kDsStageLoads = (kDsChunk + 1024 + kDsThreads - 1) / kDsThreads;

// One downsampling step: f = mf * 2^ef exactly (the float64 factor split
// into integer significand and exponent), inv_f ~ 1/f for index estimates.
struct DsStepDev {
    uint64_t mf;
    int32_t ef;
    float inv_f;
    int64_t n;
    int64_t out_offset;
};

// One series of the packed multi-step downsampling launch.
struct DsSeriesDev {
    const float* x;
    float* out;
    int64_t N;
    int32_t step0;       // first step in the step table
    int32_t nsteps;
    int32_t chunk_offset;  // first workgroup of this series
    int32_t pad;
};

// One series of the batched preparation (dereddening + normalisation).
struct PrepSeriesDev {
    const float* x;   // input series
    float* out;       // prepared series
    float* lores;     // N/sf block means (sf > 1)
    float* rmed;      // running median of lores (sf > 1) or of x (sf == 1)
    float* stats;     // 2*kNormaliseParts partials, then mean and std pairs
    int64_t N;
    int64_t nl;       // N / sf
    int64_t sf;       // scrunch factor
    int32_t width;    // running median width
    int32_t dereddened;  // normalise `out` (dereddened) rather than `x`
};

// Splits a positive float64 into (mf, ef) with f == mf * 2^ef, 2^52 <= mf < 2^53.
void split_double(double f, uint64_t* mf, int* ef);

// Launchers (all asynchronous on `stream`).
hipError_t launch_downsample_direct(const float* in, int64_t N, double f, int64_t n, float* out,
                                    hipStream_t stream);
hipError_t launch_downsample_packed(const DsSeriesDev* d_series, int nseries,
                                    const DsStepDev* d_steps, int total_chunks, int chunk,
                                    int halo, hipStream_t stream);
hipError_t launch_running_median(const float* x, int64_t N, int width, float* out,
                                 hipStream_t stream);
hipError_t launch_scrunch_mean(const float* x, int64_t nblocks, int64_t factor, float* out,
                               hipStream_t stream);
hipError_t launch_deredden_subtract(const float* x, int64_t N, const float* fp, int64_t nl,
                                    int64_t sf, float* out, hipStream_t stream);
// Dereddening and/or normalisation of nseries series in a few launches
// (series = blockIdx.y); same results as the one-series launches.
hipError_t launch_prep_batch(const PrepSeriesDev* d_ps, int nseries, int64_t max_n,
                             int64_t max_lores, int max_width, int deredden, int normalise,
                             hipStream_t stream);
// mean/std workspace: 2*(parts + 2) floats
hipError_t launch_normalise(const float* x, int64_t N, float* out, float* work, int parts,
                            hipStream_t stream);
hipError_t launch_rollback(const float* x, int64_t size, int64_t shift, float* out,
                           hipStream_t stream);
hipError_t launch_fused_rollback_add(const float* x, const float* y, int64_t size, int64_t shift,
                                     float* z, hipStream_t stream);
hipError_t launch_circular_prefix_sum(const float* x, int64_t size, int64_t nsum, float* out,
                                      hipStream_t stream);
hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t stream);
// 8-bit samples to float32 (kind 1: unsigned, 2: signed).
hipError_t launch_convert_samples(const uint8_t* raw, int64_t n, int kind, float* out,
                                  hipStream_t stream);
// Folding: in-place float32 scaling, the sum of the rows of a rows x bins
// array (rows added in order) and the downsampling of its columns.
hipError_t launch_fold_scale(float* x, int64_t count, float scale, hipStream_t stream);
hipError_t launch_fold_sum_rows(const float* x, int64_t rows, int64_t bins, float* out,
                                hipStream_t stream);
hipError_t launch_downsample_vertical(const float* x, int64_t rows, int64_t bins, double f,
                                      int64_t n, float* out, hipStream_t stream);

constexpr int kNormaliseParts = 1024;

}  // namespace rt
