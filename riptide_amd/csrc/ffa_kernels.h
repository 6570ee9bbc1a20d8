// Host launchers of the FFA / S/N kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "tasks.h"

namespace rt {

// Runs ntasks FfaTask work items (one workgroup each) with the kernels of
// `variant`; every task's ffa_variant(bins) must be `variant`.  *d_err gets
// bit 0 set if a task's cone would overrun LDS (never, for tasks from
// ffa_plan), bit 1 if a task's row length is not the variant's.
hipError_t launch_ffa_tasks(const FfaTask* d_tasks, int ntasks, int* d_err, hipStream_t stream,
                            int shape, int variant);

// Runs a task list whose host copy is h_tasks: one launch per run of tasks of
// equal kernel variant (sort_by_variant() makes the runs long).
hipError_t launch_ffa_task_list(const FfaTask* d_tasks, const FfaTask* h_tasks, int ntasks,
                                int* d_err, hipStream_t stream, int shape);

// Stable sort of independent tasks (one pass) by kernel variant.
void sort_by_variant(std::vector<FfaTask>& tasks);

// Persistent batch kernel over whole transforms of one kernel variant:
// `slots` workgroups, each with slot_floats floats of scratch (two
// intermediate states); *d_queue must be 0.
hipError_t launch_ffa_batch(const FfaTransform* d_trs, int ntr, int* d_queue, float* d_scratch,
                            int64_t slot_floats, int slots, int* d_err, hipStream_t stream,
                            int shape, int variant);

// Builds n tile records (FfaRecordSpec -> FfaRecord layout in `out`): the
// cone and row tables a task of each spec would compute itself.
hipError_t launch_ffa_records(const FfaRecordSpec* d_specs, int n, int32_t* d_out,
                              hipStream_t stream);

// Launchers of one variant (ffa_task.h, built by ffa_variant.hip).
template <int V>
hipError_t launch_ffa_tasks_variant(const FfaTask* d_tasks, int ntasks, int* d_err,
                                    hipStream_t stream, int shape);
template <int V>
hipError_t launch_ffa_batch_variant(const FfaTransform* d_trs, int ntr, int* d_queue,
                                    float* d_scratch, int64_t slot_floats, int slots, int* d_err,
                                    hipStream_t stream, int shape);

// Boxcar S/N of rows x cols profiles; long_scratch (rows*cols floats) is only
// used for cols > 1024.
hipError_t launch_snr_rows(const float* x, int64_t rows, int cols, const int32_t* widths,
                           const float* hb, int nw, float stdnoise, float* out,
                           float* long_scratch, hipStream_t stream);

// One merge level (tree depth `depth`) of a rows x cols FFA straight from HBM.
hipError_t launch_ffa_level_global(const float* src, float* dst, int64_t rows, int64_t cols,
                                   int depth, hipStream_t stream);

}  // namespace rt
