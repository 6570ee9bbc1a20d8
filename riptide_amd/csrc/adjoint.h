// Launchers of the adjoint kernels (see adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

// Largest number of trial widths of the S/N backward kernel.
constexpr int kSnrBwdMaxW = 64;

// gx[rows x bins] = d(sum_w gsnr[r,w] * snr[r,w]) / d x[r, :]
hipError_t launch_snr_backward(const float* x, int rows, int bins, const float* gsnr,
                               const int32_t* widths, const float* hb, int nw, float stdnoise,
                               float* gx, hipStream_t s);

// Transpose of the FFA merge level at tree depth `depth`.
hipError_t launch_ffa_adjoint_level(const float* gp, float* gc, int64_t rows, int64_t cols,
                                    int depth, hipStream_t s);

// Transpose of downsampling by f: gin[N] (+)= D^T gout[n]; idx is n int4 of
// workspace (the index table of the downsampling).
hipError_t launch_downsample_adjoint(const float* gout, int64_t n, double f, int64_t N,
                                     float* gin, int accumulate, int4* idx, hipStream_t s);

// flags[t] = any non-zero S/N gradient in rows [rows[t].x, rows[t].x + rows[t].y) of gsnr
// (nw per row).
hipError_t launch_transform_need(const float* gsnr, const int2* rows, int ntr, int nw, int* flags,
                                 hipStream_t s);

// gin[i] (+)= gout[i] for i < n, 0 beyond.
hipError_t launch_add_prefix(const float* gout, int64_t n, int64_t N, float* gin, int accumulate,
                             hipStream_t s);

// Transpose of y = (x - mean)/std given y and 1/std; part holds 4*nparts floats.
hipError_t launch_normalise_adjoint(const float* g, const float* y, int64_t N, float inv_std,
                                    float* part, int nparts, float* gx, hipStream_t s);

// ---- batched backward of series sharing one search plan (grad_batch.hip)

// One (series, transform) pair: its recomputed forward rows T, the S/N
// gradient A, the ping-pong buffers of the transposed merge levels B and C
// (each rows x bins), its rows of the S/N gradient and (h + b, b) table.
struct GradPair {
    const float* T;
    float* A;
    float* B;
    float* C;
    const float* gsnr;
    const float* hb;
    int32_t rows, bins, rows_eval, H;
    float stdnoise;
    int32_t pad;
};

// One downsampling step used by the batch: factor f = mf * 2^ef, its index
// table (idx; null when f == 1) of n outputs, 1/f in float32.
struct GradStep {
    uint64_t mf;
    int32_t ef;
    float inv_f;
    int64_t n;
    int4* idx;
};

// One (series, step) group: the series' normalised input x, its downsampled
// values ds (f > 1), its fold gradient gds (n floats) and its pairs [p0, p1).
struct GradGroup {
    const float* x;
    float* ds;
    float* gds;
    int32_t step;
    int32_t p0, p1;
    int32_t pad;
};

// Waves (rows) per block of the batched S/N backward.
constexpr int kGradSnrWaves = 4;

hipError_t launch_flags_batch(const float* gsnr, int64_t series_stride, const int2* rows, int ntr,
                              int nw, int nseries, int* flags, hipStream_t s);
hipError_t launch_ds_index_steps(const GradStep* steps, const int4* blk, int nblk, int64_t N,
                                 hipStream_t s);
hipError_t launch_ds_values_groups(const GradGroup* groups, const GradStep* steps,
                                   const int4* blk, int nblk, hipStream_t s);
hipError_t launch_fill_pairs(const GradPair* pairs, const int4* blk, int nblk, hipStream_t s);
hipError_t launch_snr_backward_pairs(const GradPair* pairs, const int4* blk, int nblk,
                                     int max_bins, const int32_t* widths, int nw, hipStream_t s);
hipError_t launch_ffa_adjoint_level_pairs(const GradPair* pairs, const int4* blk, int nblk,
                                          hipStream_t s);
hipError_t launch_fold_groups(const GradGroup* groups, const GradPair* pairs,
                              const GradStep* steps, const int4* blk, int nblk, hipStream_t s);
hipError_t launch_ds_adjoint_series(const GradGroup* groups, const GradStep* steps, const int* g0,
                                    int nseries, int64_t N, float* gz, hipStream_t s);
hipError_t launch_normalise_adjoint_batch(const float* g, const float* y, int64_t N, int nseries,
                                          const float* inv_std, float* part, int nparts,
                                          float* gx, hipStream_t s);

}  // namespace rt
