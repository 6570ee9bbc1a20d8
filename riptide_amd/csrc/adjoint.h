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

}  // namespace rt
