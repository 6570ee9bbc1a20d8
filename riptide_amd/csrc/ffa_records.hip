// Tile records of the batched search: the cones and row tables that the
// tasks of transforms shared by several series would otherwise each compute
// (FfaRecord layout in tasks.h), built once per run by the same device
// functions the tasks use.
#include <hip/hip_runtime.h>

#include "ffa_kernels.h"
#include "ffa_task.h"
#include "tasks.h"

namespace rt {

// One wavefront per tile record: the merge coefficients, the cone (cone
// tasks) and the row tables of every level, computed by the same functions a
// task uses (ffa_task.h) with the node's first row at 0, written to global
// memory.
__global__ __launch_bounds__(64) void ffa_record_kernel(const FfaRecordSpec* __restrict__ specs,
                                                        int n, int32_t* __restrict__ out)
{
    __shared__ Seg seg[2 << kMaxConeLevels];
    __shared__ MergeCoef coef;
    __shared__ int lrows[kMaxConeLevels + 1];
    const int r = blockIdx.x;
    if (r >= n) return;
    const FfaRecordSpec sp = specs[r];
    const int lane = threadIdx.x;
    int32_t* rec = out + sp.off;
    FfaTask tk = {};
    tk.bins = sp.bins;
    tk.node_r0 = 0;
    tk.node_m = sp.node_m;
    tk.lo = sp.lo;
    tk.hi = sp.hi;
    tk.levels = sp.levels;
    tk.cone = sp.cone;
    fill_merge_coef(&coef, sp.node_m, sp.levels, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t* tab = reinterpret_cast<uint32_t*>(rec + kRecHeader + sp.cap);
    if (sp.cone) {
        build_cone_wave(seg, tk, &coef, lrows, lane);
        const int bottom = sp.levels > 0 ? lrows[sp.levels] : sp.hi - sp.lo;
        if (lane == 0) rec[0] = bottom;
        if (lane <= sp.levels) rec[1 + lane] = lrows[lane];
        for (int p = lane; p < bottom; p += kWave) {
            const Seg g = seg[seg_find(seg, sp.levels, p)];
            rec[kRecHeader + p] = g.r0 + g.lo + (p - g.base);
        }
        for (int e = lane; e < sp.levels * sp.cap; e += kWave) {
            const int t = e / sp.cap;
            const int p = e - t * sp.cap;
            if (p < lrows[t]) {
                int a, b, sh;
                cone_row_map(seg, t, p, sp.bins, sp.node_m, &coef, &a, &b, &sh);
                tab[e] = pack_row(a, b, sh);
            }
        }
    } else {
        if (lane == 0) rec[0] = sp.node_m;
        for (int p = lane; p < sp.node_m; p += kWave)
            identity_row_tables(sp.node_m, sp.levels, p, sp.bins, &coef, tab, sp.cap);
    }
}

hipError_t launch_ffa_records(const FfaRecordSpec* d_specs, int n, int32_t* d_out,
                              hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ffa_record_kernel, dim3(n), dim3(kWave), 0, stream, d_specs, n, d_out);
    return hipGetLastError();
}

}  // namespace rt
