// One variant of the FFA task kernels: compiled once per variant index
// (-DRT_FFA_VARIANT=v, v = 0 .. kFfaVariants-1), so that every translation
// unit holds the merge code of its own row lengths only.
#include "ffa_task.h"

#ifndef RT_FFA_VARIANT
#error "RT_FFA_VARIANT must be defined"
#endif

namespace rt {

static_assert(RT_FFA_VARIANT >= 0 && RT_FFA_VARIANT < kFfaVariants, "unknown FFA kernel variant");

template hipError_t launch_ffa_tasks_variant<RT_FFA_VARIANT>(const FfaTask*, int, int*,
                                                             hipStream_t, int);
template hipError_t launch_ffa_batch_variant<RT_FFA_VARIANT>(const FfaTransform*, int, int*,
                                                             float*, int64_t, int, int*,
                                                             hipStream_t, int);

}  // namespace rt

#ifdef RT_PHASE_TIMING
#define RT_PHASE_CAT2(a, b) a##b
#define RT_PHASE_CAT(a, b) RT_PHASE_CAT2(a, b)
// Phase totals of this variant's batch kernel since the last call (cycles,
// kPhases values), then cleared.
extern "C" int RT_PHASE_CAT(rt_phase_read_v, RT_FFA_VARIANT)(unsigned long long* out)
{
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(rt::g_rt_phase), sizeof(rt::g_rt_phase)) != hipSuccess)
        return 1;
    static const unsigned long long zero[rt::kPhases] = {};
    return hipMemcpyToSymbol(HIP_SYMBOL(rt::g_rt_phase), zero, sizeof(zero)) != hipSuccess;
}
#endif
