// Host launchers of the FFA / S/N kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "tasks.h"

namespace rt {

// Runs ntasks FfaTask work items (one workgroup each).  *d_err gets bit 0 set
// if a task's cone would overrun LDS (never, for tasks from ffa_plan).
hipError_t launch_ffa_tasks(const FfaTask* d_tasks, int ntasks, int* d_err, hipStream_t stream,
                            int shape);

// Persistent batch kernel over whole transforms: `slots` workgroups, each with
// slot_floats floats of scratch (two intermediate states); *d_queue must be 0.
hipError_t launch_ffa_batch(const FfaTransform* d_trs, int ntr, int* d_queue, float* d_scratch,
                            int64_t slot_floats, int slots, int* d_err, hipStream_t stream,
                            int shape);

// Boxcar S/N of rows x cols profiles; long_scratch (rows*cols floats) is only
// used for cols > 1024.
hipError_t launch_snr_rows(const float* x, int64_t rows, int cols, const int32_t* widths,
                           const float* hb, int nw, float stdnoise, float* out,
                           float* long_scratch, hipStream_t stream);

// One merge level (tree depth `depth`) of a rows x cols FFA straight from HBM.
hipError_t launch_ffa_level_global(const float* src, float* dst, int64_t rows, int64_t cols,
                                   int depth, hipStream_t stream);

}  // namespace rt
