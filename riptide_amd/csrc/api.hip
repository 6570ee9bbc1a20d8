// C-ABI entry points (include/riptide_amd.h): argument checks with the
// reference's messages, the device scratch pool, and the launch schedules.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/riptide_amd.h"
#include "adjoint.h"
#include "exact_f64.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "plan.h"
#include "series_kernels.h"
#include "tasks.h"

namespace rt {

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

void set_error(const std::string& msg) { g_err = msg; }

// Workgroup shape of the FFA kernels for calls made from now on.
static int g_ffa_shape = -1;
int current_ffa_shape() { return g_ffa_shape < 0 ? ffa_default_shape() : g_ffa_shape; }

#define RT_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(RT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Device scratch pool: grow-only buffers per (device, slot).  Calls that use
// it hold g_pool_mu, so one device runs one such call at a time.
// ---------------------------------------------------------------------------

enum Slot {
    kSlotTasks = 0,
    kSlotTables,
    kSlotDs,
    kSlotInter,
    kSlotMisc,
    kSlotErr,
    kSlotCount
};

struct PoolBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

static std::mutex g_pool_mu;
static std::map<int, std::vector<PoolBuf>> g_pool;

static int pool_get(int slot, size_t bytes, void** out)
{
    int dev = 0;
    RT_HIP(hipGetDevice(&dev));
    auto& v = g_pool[dev];
    if (v.empty()) v.resize(kSlotCount);
    PoolBuf& b = v[slot];
    if (b.bytes < bytes) {
        if (b.p) {
            RT_HIP(hipDeviceSynchronize());
            RT_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes, 256);
        RT_HIP(hipMalloc(&b.p, want));
        b.bytes = want;
    }
    *out = b.p;
    return RT_OK;
}

// Regions carved in order, 256-byte aligned, from one device buffer that the
// caller owns (the allocation-free entry point's static array): the
// alternative to the pool for calls that must not allocate.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
};

static int arena_get(Arena* a, size_t bytes, void** out)
{
    const size_t off = (a->used + 255) & ~(size_t)255;
    if (off > a->cap || bytes > a->cap - off)
        return fail(RT_EINVAL, "periodogram needs more scratch than the allocation-free entry "
                               "point holds (" + std::to_string(a->cap >> 20) + " MiB)");
    *out = a->base + off;
    a->used = off + bytes;
    return RT_OK;
}

// A buffer of a pool slot, or the next region of `arena` when one is given.
static int scratch_get(Arena* arena, int slot, size_t bytes, void** out)
{
    return arena ? arena_get(arena, bytes, out) : pool_get(slot, bytes, out);
}

template <typename T>
static int upload(int slot, const std::vector<T>& v, T** d, hipStream_t stream,
                  Arena* arena = nullptr)
{
    void* p = nullptr;
    int st = scratch_get(arena, slot, v.size() * sizeof(T), &p);
    if (st) return st;
    if (!v.empty())
        RT_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    *d = static_cast<T*>(p);
    return RT_OK;
}

// ---------------------------------------------------------------------------
// Argument checks (messages of the reference)
// ---------------------------------------------------------------------------

static int check_widths(const uint64_t* w, int64_t nw, int64_t bins)
{
    // snr.hpp:21-31
    for (int64_t i = 0; i < nw; ++i)
        if (!(w[i] > 0 && w[i] < (uint64_t)bins))
            return fail(RT_EINVAL, "trial widths must be all > 0 and < columns");
    return RT_OK;
}

static int check_stdnoise(float s)
{
    // snr.hpp:14-18
    if (!(s > 0)) return fail(RT_EINVAL, "stdnoise must be > 0");
    return RT_OK;
}

// (h + b, b) of every width, for bins in [b0, b0 + nb)
static std::vector<float> hb_table(int64_t b0, int64_t nb, const uint64_t* w, int64_t nw)
{
    std::vector<float> t((size_t)(nb * nw * 2));
    for (int64_t b = 0; b < nb; ++b)
        for (int64_t i = 0; i < nw; ++i) {
            float hpb = 0, bb = 0;
            if (w[i] < (uint64_t)(b0 + b)) boxcar_hb(b0 + b, (int64_t)w[i], &hpb, &bb);
            t[(size_t)((b * nw + i) * 2)] = hpb;
            t[(size_t)((b * nw + i) * 2 + 1)] = bb;
        }
    return t;
}

// ---------------------------------------------------------------------------
// FFA task execution
// ---------------------------------------------------------------------------

// Uploads all passes' tasks, launches them in order, waits and checks the
// kernel's overflow flag.
static int run_passes(const std::vector<std::vector<FfaTask>>& passes, hipStream_t stream,
                      bool sync, int shape, Arena* arena = nullptr)
{
    // the tasks of a pass are independent: grouped by kernel variant, one
    // launch per variant
    std::vector<FfaTask> all;
    std::vector<size_t> first;
    for (const auto& p : passes) {
        first.push_back(all.size());
        const size_t at = all.size();
        all.insert(all.end(), p.begin(), p.end());
        // most passes hold one variant only; sort the others in place
        bool mixed = false;
        for (size_t i = 1; i < p.size() && !mixed; ++i)
            mixed = ffa_variant(p[i].bins) != ffa_variant(p[0].bins);
        if (mixed) {
            std::vector<FfaTask> q(all.begin() + at, all.end());
            sort_by_variant(q);
            std::copy(q.begin(), q.end(), all.begin() + at);
        }
    }
    if (all.empty()) return RT_OK;
    FfaTask* d_tasks = nullptr;
    int st = upload(kSlotTasks, all, &d_tasks, stream, arena);
    if (st) return st;
    void* errp = nullptr;
    st = scratch_get(arena, kSlotErr, sizeof(int), &errp);
    if (st) return st;
    RT_HIP(hipMemsetAsync(errp, 0, sizeof(int), stream));
    for (size_t i = 0; i < passes.size(); ++i)
        RT_HIP(launch_ffa_task_list(d_tasks + first[i], all.data() + first[i],
                                    (int)passes[i].size(), (int*)errp, stream, shape));
    if (sync) {
        int herr = 0;
        RT_HIP(hipMemcpyAsync(&herr, errp, sizeof(int), hipMemcpyDeviceToHost, stream));
        RT_HIP(hipStreamSynchronize(stream));
        if (herr & 1) return fail(RT_EDEVICE, "FFA tile exceeded its LDS budget");
        if (herr) return fail(RT_EDEVICE, "FFA task launched with the wrong kernel variant");
    }
    return RT_OK;
}

// Global-memory FFA for blocks whose rows are too wide to tile in LDS.
static int ffa_global(const float* d_in, int64_t rows, int64_t cols, float* d_out,
                      hipStream_t stream)
{
    const int H = ffa_height(rows);
    if (H == 0) {
        RT_HIP(hipMemcpyAsync(d_out, d_in, (size_t)(rows * cols) * sizeof(float),
                              hipMemcpyDeviceToDevice, stream));
        return RT_OK;
    }
    void* tmp = nullptr;
    int st = pool_get(kSlotInter, (size_t)(rows * cols) * sizeof(float), &tmp);
    if (st) return st;
    // depth H-1 first; the last (depth 0) launch must write d_out
    const float* src = d_in;
    for (int d = H - 1; d >= 0; --d) {
        const bool to_out = ((H - 1 - d) % 2) == ((H - 1) % 2);
        float* dst = to_out ? d_out : (float*)tmp;
        RT_HIP(launch_ffa_level_global(src, dst, rows, cols, d, stream));
        src = dst;
    }
    return RT_OK;
}

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Minimum rows per workgroup for the LDS path.
static constexpr int kMinCapRows = 16;

}  // namespace rt

using namespace rt;

extern "C" {

const char* rt_last_error(void) { return g_err.c_str(); }

int rt_version(void) { return 1; }

int rt_set_ffa_shape(int shape)
{
    if (shape < -1 || shape >= kFfaShapes) return fail(RT_EINVAL, "unknown FFA workgroup shape");
    g_ffa_shape = shape;
    return RT_OK;
}

int rt_periodogram_length(int64_t size, double tsamp, double pmin, double pmax, int64_t bmin,
                          int64_t bmax, int64_t* length)
{
    SeriesPlan p;
    const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &p);
    if (!e.empty()) return fail(RT_EINVAL, e);
    *length = p.length;
    return RT_OK;
}

int rt_periodogram_grid(int64_t size, double tsamp, double pmin, double pmax, int64_t bmin,
                        int64_t bmax, double* h_periods, uint32_t* h_foldbins)
{
    SeriesPlan p;
    const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &p);
    if (!e.empty()) return fail(RT_EINVAL, e);
    plan_periods(p, h_periods, h_foldbins);
    return RT_OK;
}

int rt_downsample_index(int64_t k, double f, int64_t size, int64_t* imin, int64_t* imax,
                        float* wmin, float* wmax)
{
    if (!(f > 0)) return fail(RT_EINVAL, "f must be > 0");
    uint64_t mf;
    int ef;
    split_double(f, &mf, &ef);
    const DsIndex ix = ds_index_fast(k, mf, ef, size);
    *imin = ix.imin;
    *imax = ix.imax;
    *wmin = ix.wmin;
    *wmax = ix.wmax;
    return RT_OK;
}

int rt_ffa2(const float* d_in, int64_t rows, int64_t cols, float* d_out, void* stream)
{
    if (rows < 1 || cols < 1)
        return fail(RT_EINVAL, "ffa2 input must have at least one row and one column");
    if (rows > (1LL << 30) || rows * cols > (1LL << 40))
        return fail(RT_EINVAL, "ffa2 input too large");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    const int shape = current_ffa_shape();
    const int cap = ffa_cap_rows((int)std::min<int64_t>(cols, 1 << 20), shape);
    if (cols > (1 << 20) || cap < kMinCapRows || rows > (1LL << 31) / std::max<int64_t>(cols, 1)) {
        const int st = ffa_global(d_in, rows, cols, d_out, s);
        if (st) return st;
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    TransformSpec sp;
    sp.m = rows;
    sp.bins = (int)cols;
    sp.cap = cap;
    sp.shape = shape;
    sp.src = d_in;
    sp.kind = 0;
    sp.dst = d_out;
    if (rows > cap) {
        void* inter = nullptr;
        const size_t n = (size_t)(rows * cols);
        int st = pool_get(kSlotInter, 2 * n * sizeof(float), &inter);
        if (st) return st;
        sp.buf_a = (float*)inter;
        sp.buf_b = (float*)inter + n;
    }
    std::vector<std::vector<FfaTask>> passes;
    if (plan_transform_tasks(sp, passes) < 0) {
        const int st = ffa_global(d_in, rows, cols, d_out, s);
        if (st) return st;
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    return run_passes(passes, s, true, shape);
}

int rt_time_ffa(int64_t rows, int64_t cols, int shape, int with_snr, int64_t reps, double* ms)
{
    // Device time of one FFA transform (+ S/N of every row, 10 widths) of a
    // zero-filled rows x cols block with workgroup shape `shape`: planning and
    // uploads happen once, then the passes are launched `reps` times between
    // two events.
    if (rows < 2 || cols < 2 || reps < 1) return fail(RT_EINVAL, "rows, cols >= 2 and reps >= 1");
    if (shape < 0 || shape >= kFfaShapes) return fail(RT_EINVAL, "unknown FFA workgroup shape");
    const int cap = ffa_cap_rows((int)cols, shape);
    if (cap < kMinCapRows) return fail(RT_EINVAL, "rows too wide for the LDS path");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    const size_t n = (size_t)(rows * cols);
    float *in = nullptr, *out = nullptr, *inter = nullptr, *tab = nullptr;
    int* err = nullptr;
    FfaTask* d_tasks = nullptr;
    RT_HIP(hipMalloc(&in, n * sizeof(float)));
    RT_HIP(hipMalloc(&out, std::max<size_t>(n, (size_t)rows * 16) * sizeof(float)));
    RT_HIP(hipMalloc(&inter, 2 * n * sizeof(float)));
    RT_HIP(hipMemset(in, 0, n * sizeof(float)));
    std::vector<uint64_t> w = {1, 2, 3, 4, 6, 9, 13, 19, 28, 42};
    while (!w.empty() && w.back() >= (uint64_t)cols) w.pop_back();
    std::vector<float> t = hb_table(cols, 1, w.data(), (int64_t)w.size());
    const size_t woff = t.size();
    t.resize(woff + w.size());
    for (size_t i = 0; i < w.size(); ++i) {
        const int32_t v = (int32_t)w[i];
        std::memcpy(&t[woff + i], &v, sizeof(int32_t));
    }
    RT_HIP(hipMalloc(&tab, t.size() * sizeof(float)));
    RT_HIP(hipMemcpy(tab, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    RT_HIP(hipMalloc(&err, sizeof(int)));
    RT_HIP(hipMemset(err, 0, sizeof(int)));
    TransformSpec sp;
    sp.m = rows;
    sp.bins = (int)cols;
    sp.cap = cap;
    sp.shape = shape;
    sp.src = in;
    sp.buf_a = inter;
    sp.buf_b = inter + n;
    sp.kind = with_snr ? 1 : 0;
    sp.dst = out;
    sp.snr = out;
    sp.widths = (const int32_t*)(tab + woff);
    sp.hb = tab;
    sp.nw = (int)w.size();
    sp.stdnoise = 1.0f;
    sp.rows_eval = (int)rows;
    sp.wmax = w.empty() ? 0 : (int)w.back();
    std::vector<std::vector<FfaTask>> passes;
    if (plan_transform_tasks(sp, passes) < 0) return fail(RT_EINVAL, "cannot tile this block");
    std::vector<FfaTask> all;
    std::vector<size_t> first;
    for (auto& q : passes) {
        first.push_back(all.size());
        all.insert(all.end(), q.begin(), q.end());
    }
    RT_HIP(hipMalloc(&d_tasks, all.size() * sizeof(FfaTask)));
    RT_HIP(hipMemcpy(d_tasks, all.data(), all.size() * sizeof(FfaTask), hipMemcpyHostToDevice));
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    for (int64_t r = -1; r < reps; ++r) {
        if (r == 0) RT_HIP(hipEventRecord(a, nullptr));
        for (size_t i = 0; i < passes.size(); ++i)
            RT_HIP(launch_ffa_task_list(d_tasks + first[i], all.data() + first[i],
                                        (int)passes[i].size(), err, nullptr, shape));
    }
    RT_HIP(hipEventRecord(b, nullptr));
    RT_HIP(hipEventSynchronize(b));
    float t_ms = 0;
    RT_HIP(hipEventElapsedTime(&t_ms, a, b));
    *ms = t_ms / (double)reps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    for (void* q : {(void*)in, (void*)out, (void*)inter, (void*)tab, (void*)err, (void*)d_tasks})
        hipFree(q);
    return RT_OK;
}

int rt_snr2(const float* d_in, int64_t rows, int64_t cols, const uint64_t* h_widths, int64_t nw,
            float stdnoise, float* d_out, void* stream)
{
    // python_bindings.cpp:131-132: stdnoise is checked before the widths
    int st = check_stdnoise(stdnoise);
    if (st) return st;
    st = check_widths(h_widths, nw, cols);
    if (st) return st;
    if (rows == 0 || nw == 0) return RT_OK;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    std::vector<int32_t> w32(h_widths, h_widths + nw);
    std::vector<float> hb = hb_table(cols, 1, h_widths, nw);
    std::vector<float> tab(hb);
    const size_t woff = tab.size();
    tab.resize(woff + (size_t)nw);
    std::memcpy(tab.data() + woff, w32.data(), (size_t)nw * sizeof(int32_t));
    float* d_tab = nullptr;
    st = upload(kSlotTables, tab, &d_tab, s);
    if (st) return st;
    float* scratch = nullptr;
    if (cols > kWave * kSnrRegsMaxK) {
        void* p = nullptr;
        st = pool_get(kSlotInter, (size_t)(rows * cols) * sizeof(float), &p);
        if (st) return st;
        scratch = (float*)p;
    }
    RT_HIP(launch_snr_rows(d_in, rows, (int)cols, (const int32_t*)(d_tab + woff), d_tab, (int)nw,
                           stdnoise, d_out, scratch, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_downsample(const float* d_in, int64_t size, double f, float* d_out, void* stream)
{
    // downsample.hpp:11-16
    if (!((f > 1.0) & (f <= (double)(size_t)size)))
        return fail(RT_EINVAL, "Downsampling factor must verify: 1 < f <= size");
    hipStream_t s = S(stream);
    const int64_t n = downsampled_size(size, f);
    RT_HIP(launch_downsample_direct(d_in, size, f, n, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_convert_samples(const void* d_raw, int64_t n, int nbits, int is_signed, float* d_out,
                       void* stream)
{
    // time_series.py:367-375: 8-bit samples are cast to float32
    if (nbits != 8) return fail(RT_EINVAL, "only 8-bit samples need converting");
    if (n < 1) return RT_OK;
    RT_HIP(launch_convert_samples((const uint8_t*)d_raw, n, is_signed ? 2 : 1, d_out, S(stream)));
    return RT_OK;
}

int rt_fold_rows(float* d_x, int64_t rows, int64_t bins, float scale, int sum, float* d_out,
                 void* stream)
{
    if (rows < 1 || bins < 1) return fail(RT_EINVAL, "nothing to fold");
    hipStream_t s = S(stream);
    RT_HIP(launch_fold_scale(d_x, rows * bins, scale, s));
    if (sum) RT_HIP(launch_fold_sum_rows(d_x, rows, bins, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_downsample_vertical(const float* d_x, int64_t rows, int64_t bins, double f, float* d_out,
                           void* stream)
{
    // folding.py:9-14 (the checks of downsample_vertical) and downsample.hpp:11-16
    if (!(f > 1.0)) return fail(RT_EINVAL, "factor must be > 1");
    if (!(f < (double)rows))
        return fail(RT_EINVAL, "factor must be strictly smaller than the number of input lines");
    hipStream_t s = S(stream);
    const int64_t n = downsampled_size(rows, f);
    RT_HIP(launch_downsample_vertical(d_x, rows, bins, f, n, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_running_median(const float* d_in, int64_t size, int64_t width, float* d_out, void* stream)
{
    // running_median.hpp:103-107
    if (!(width % 2)) return fail(RT_EINVAL, "width must be an odd number");
    if (!(width < size)) return fail(RT_EINVAL, "width must be < size");
    hipStream_t s = S(stream);
    RT_HIP(launch_running_median(d_in, size, (int)width, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_rollback(const float* d_in, int64_t size, int64_t shift, float* d_out, void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "size must be >= 1");
    if (shift < 0) return fail(RT_EINVAL, "shift must be positive");
    hipStream_t s = S(stream);
    RT_HIP(launch_rollback(d_in, size, shift, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_fused_rollback_add(const float* d_x, const float* d_y, int64_t size, int64_t shift,
                          float* d_out, void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "size must be >= 1");
    if (shift < 0) return fail(RT_EINVAL, "shift must be positive");
    hipStream_t s = S(stream);
    RT_HIP(launch_fused_rollback_add(d_x, d_y, size, shift, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_circular_prefix_sum(const float* d_in, int64_t size, int64_t nsum, float* d_out,
                           void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "size must be >= 1");
    if (nsum < 0) return fail(RT_EINVAL, "nsum must be >= 0");
    if (nsum == 0) return RT_OK;
    hipStream_t s = S(stream);
    RT_HIP(launch_circular_prefix_sum(d_in, size, nsum, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_benchmark_ffa2(int64_t rows, int64_t cols, int64_t loops, double* seconds)
{
    if (rows < 1 || cols < 1 || loops < 1)
        return fail(RT_EINVAL, "rows, cols and loops must be >= 1");
    float* in = nullptr;
    float* out = nullptr;
    const size_t n = (size_t)(rows * cols);
    RT_HIP(hipMalloc(&in, n * sizeof(float)));
    RT_HIP(hipMalloc(&out, n * sizeof(float)));
    RT_HIP(hipMemset(in, 0, n * sizeof(float)));
    int st = rt_ffa2(in, rows, cols, out, nullptr);  // warm-up
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    RT_HIP(hipEventRecord(a, nullptr));
    for (int64_t i = 0; i < loops && st == RT_OK; ++i)
        st = rt_ffa2(in, rows, cols, out, nullptr);
    RT_HIP(hipEventRecord(b, nullptr));
    RT_HIP(hipEventSynchronize(b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    hipFree(in);
    hipFree(out);
    if (st) return st;
    *seconds = ms * 1e-3 / (double)loops;
    return RT_OK;
}

int rt_scrunch_mean(const float* d_in, int64_t size, int64_t factor, float* d_out, void* stream)
{
    if (factor < 1) return fail(RT_EINVAL, "factor must be >= 1");
    hipStream_t s = S(stream);
    const int64_t nb = size / factor;
    if (nb > 0) RT_HIP(launch_scrunch_mean(d_in, nb, factor, d_out, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

static int fast_rmed(const float* d_in, int64_t size, int64_t width_samples, int64_t min_points,
                     float* d_out, bool subtract, hipStream_t s)
{
    // running_medians.py:73-85
    if (!(min_points % 2)) return fail(RT_EINVAL, "min_points must be an odd number");
    const int64_t sf = (int64_t)std::max(1.0, width_samples / (double)min_points);
    if (sf == 1) {
        if (!(width_samples % 2)) return fail(RT_EINVAL, "width must be an odd number");
        if (!(width_samples < size)) return fail(RT_EINVAL, "width must be < size");
        if (!subtract) {
            RT_HIP(launch_running_median(d_in, size, (int)width_samples, d_out, s));
            return RT_OK;
        }
        void* p = nullptr;
        int st = pool_get(kSlotMisc, (size_t)size * sizeof(float), &p);
        if (st) return st;
        RT_HIP(launch_running_median(d_in, size, (int)width_samples, (float*)p, s));
        // data - rmed with rmed at full resolution: the interpolation kernel
        // with sf = 1 reduces to a plain float32 subtraction
        RT_HIP(launch_deredden_subtract(d_in, size, (const float*)p, size, 1, d_out, s));
        return RT_OK;
    }
    const int64_t nl = size / sf;
    if (!(min_points < nl)) return fail(RT_EINVAL, "width must be < size");
    void* p = nullptr;
    int st = pool_get(kSlotMisc, (size_t)(2 * nl) * sizeof(float), &p);
    if (st) return st;
    float* lores = (float*)p;
    float* rmed = lores + nl;
    RT_HIP(launch_scrunch_mean(d_in, nl, sf, lores, s));
    RT_HIP(launch_running_median(lores, nl, (int)min_points, rmed, s));
    RT_HIP(launch_deredden_subtract(subtract ? d_in : nullptr, size, rmed, nl, sf, d_out, s));
    return RT_OK;
}

int rt_deredden(const float* d_in, int64_t size, int64_t width_samples, int64_t min_points,
                float* d_out, void* stream)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    int st = fast_rmed(d_in, size, width_samples, min_points, d_out, true, s);
    if (st) return st;
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_fast_running_median(const float* d_in, int64_t size, int64_t width_samples,
                           int64_t min_points, float* d_out, void* stream)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    int st = fast_rmed(d_in, size, width_samples, min_points, d_out, false, s);
    if (st) return st;
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_normalise(const float* d_in, int64_t size, float* d_out, void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "cannot normalise an empty series");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    void* p = nullptr;
    int st = pool_get(kSlotMisc, (size_t)(2 * (kNormaliseParts + 2)) * sizeof(float), &p);
    if (st) return st;
    RT_HIP(launch_normalise(d_in, size, d_out, (float*)p, kNormaliseParts, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Periodogram
// ---------------------------------------------------------------------------

namespace rt {

struct PgramSetup {
    SeriesPlan plan;
    int shape = 0;
    float* d_hb = nullptr;
    int32_t* d_w = nullptr;
    float* d_ds = nullptr;
    int64_t nw = 0;
    int wmax = 0;
};

// Plans the search, checks the widths, uploads the period grid and the S/N
// tables and runs the packed downsampling of every step.
static int pgram_setup(const float* d_in, int64_t size, double tsamp, const uint64_t* h_widths,
                       int64_t nw, double pmin, double pmax, int64_t bmin, int64_t bmax,
                       double* d_periods, uint32_t* d_foldbins, hipStream_t s, PgramSetup* out,
                       Arena* arena = nullptr)
{
    const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &out->plan);
    if (!e.empty()) return fail(RT_EINVAL, e);
    out->shape = current_ffa_shape();
    const SeriesPlan& p = out->plan;
    int st = check_widths(h_widths, nw, bmin);
    if (st) return st;
    out->nw = nw;
    out->wmax = 0;
    for (int64_t i = 0; i < nw; ++i) out->wmax = std::max<int>(out->wmax, (int)h_widths[i]);

    std::vector<double> periods((size_t)p.length);
    std::vector<uint32_t> foldbins((size_t)p.length);
    plan_periods(p, periods.data(), foldbins.data());
    if (p.length > 0) {
        RT_HIP(hipMemcpyAsync(d_periods, periods.data(), periods.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(d_foldbins, foldbins.data(), foldbins.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    }
    std::vector<float> tab = hb_table(bmin, bmax - bmin + 1, h_widths, nw);
    const size_t woff = tab.size();
    tab.resize(woff + (size_t)nw);
    for (int64_t i = 0; i < nw; ++i) {
        const int32_t w = (int32_t)h_widths[i];
        std::memcpy(&tab[woff + (size_t)i], &w, sizeof(int32_t));
    }
    float* d_tab = nullptr;
    st = upload(kSlotTables, tab, &d_tab, s, arena);
    if (st) return st;
    out->d_hb = d_tab;
    out->d_w = (int32_t*)(d_tab + woff);

    // packed downsampling of every f > 1 step
    void* dsp = nullptr;
    st = scratch_get(arena, kSlotDs, (size_t)std::max<int64_t>(p.ds_total, 1) * sizeof(float),
                     &dsp);
    if (st) return st;
    out->d_ds = (float*)dsp;
    std::vector<DsStepDev> steps;
    double fmax = 1.0;
    for (const Step& stp : p.steps) {
        if (stp.ds_offset < 0 || stp.n == 0) continue;
        if (stp.f <= kDsPackedMaxF) {
            DsStepDev d;
            split_double(stp.f, &d.mf, &d.ef);
            d.inv_f = (float)(1.0 / stp.f);
            d.n = stp.n;
            d.out_offset = stp.ds_offset;
            steps.push_back(d);
            fmax = std::max(fmax, stp.f);
        } else {
            RT_HIP(launch_downsample_direct(d_in, size, stp.f, stp.n, out->d_ds + stp.ds_offset, s));
        }
    }
    if (!steps.empty()) {
        // one launch per kDsMaxSteps steps (the packed kernel's LDS tables),
        // all with the halo of the largest factor, as the batch path stages
        // its chunks (the float-float prefix sums are cut into per-thread
        // segments of the staged length, so the halo sets their rounding)
        const int chunk = kDsChunk;
        const int halo = (int)std::ceil(fmax) + 2;
        const int nchunks = (int)((size + chunk - 1) / chunk);
        const size_t nl = (steps.size() + kDsMaxSteps - 1) / kDsMaxSteps;
        std::vector<uint8_t> blob(nl * sizeof(DsSeriesDev) + steps.size() * sizeof(DsStepDev));
        for (size_t l = 0; l < nl; ++l) {
            DsSeriesDev sd;
            sd.x = d_in;
            sd.out = out->d_ds;
            sd.N = size;
            sd.step0 = (int32_t)(l * kDsMaxSteps);
            sd.nsteps = (int32_t)std::min<size_t>(kDsMaxSteps, steps.size() - l * kDsMaxSteps);
            sd.chunk_offset = 0;
            sd.pad = 0;
            std::memcpy(blob.data() + l * sizeof(DsSeriesDev), &sd, sizeof(sd));
        }
        std::memcpy(blob.data() + nl * sizeof(DsSeriesDev), steps.data(),
                    steps.size() * sizeof(DsStepDev));
        // the descriptors ride in the task slot's neighbour: a misc upload
        uint8_t* d_blob = nullptr;
        st = upload(kSlotMisc, blob, &d_blob, s, arena);
        if (st) return st;
        const DsStepDev* d_steps = (const DsStepDev*)(d_blob + nl * sizeof(DsSeriesDev));
        for (size_t l = 0; l < nl; ++l)
            RT_HIP(launch_downsample_packed((const DsSeriesDev*)d_blob + l, 1, d_steps, nchunks,
                                            chunk, halo, s));
    }
    // the host staging vectors above die with this frame
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

// Task specification of transform `t` of step `stp` (S/N into d_snrs); the
// intermediate buffers are left to the caller.
static TransformSpec pgram_spec(const PgramSetup& su, const float* d_in, const Step& stp,
                                const Transform& t, float* d_snrs)
{
    const SeriesPlan& p = su.plan;
    TransformSpec sp;
    sp.m = t.rows;
    sp.bins = t.bins;
    sp.cap = ffa_cap_rows(t.bins, su.shape);
    sp.shape = su.shape;
    sp.src = (stp.ds_offset < 0) ? d_in : su.d_ds + stp.ds_offset;
    sp.kind = 1;
    sp.wmax = su.wmax;
    sp.snr = d_snrs + t.out_row * su.nw;
    sp.widths = su.d_w;
    sp.hb = su.d_hb + (size_t)(t.bins - p.bmin) * su.nw * 2;
    sp.nw = (int)su.nw;
    sp.stdnoise = t.stdnoise;
    sp.rows_eval = t.rows_eval;
    return sp;
}

// Builds the tasks of every transform of steps [s0, s1) into `passes`, with
// intermediates carved from `inter` (returns floats used).
static int64_t pgram_tasks(const PgramSetup& su, const float* d_in, int s0, int s1, float* d_snrs,
                           float* inter, std::vector<std::vector<FfaTask>>& passes, bool dry)
{
    const SeriesPlan& p = su.plan;
    int64_t used = 0;
    for (int is = s0; is < s1; ++is) {
        const Step& stp = p.steps[is];
        for (int it = stp.first_tr; it < stp.first_tr + stp.ntr; ++it) {
            const Transform& t = p.tr[it];
            if (t.rows_eval <= 0) continue;
            const int64_t n = (int64_t)t.rows * t.bins;
            TransformSpec sp = pgram_spec(su, d_in, stp, t, d_snrs);
            if (t.rows > sp.cap) {
                if (!dry) {
                    sp.buf_a = inter + used;
                    sp.buf_b = inter + used + n;
                }
                used += 2 * n;
            }
            if (!dry && plan_transform_tasks(sp, passes) < 0) return -1;
        }
    }
    return used;
}

std::vector<int32_t> batch_tiling_flat();

// A one-series search runs as a batch of one series: the persistent batch
// kernel, with the batch's task lists and tile records built once.  The
// batches of the last few searches are kept, with their period and fold-bins
// grids in device memory, and rebound to each call's input and output while
// a search (device, size, sampling time, widths, period and bins ranges, FFA
// shape and batch tiling table) repeats -- as it does when every DM trial is
// searched over the same few period ranges.
struct OneSeriesBatch {
    void* handle = nullptr;
    int device = -1;
    int64_t size = 0, bmin = 0, bmax = 0;
    double tsamp = 0.0, pmin = 0.0, pmax = 0.0;
    int shape = -1;
    std::vector<uint64_t> widths;
    std::vector<int32_t> tiling;
    int64_t length = 0;
    double* d_periods = nullptr;
    uint32_t* d_foldbins = nullptr;
};
// searches kept (most recently used first)
constexpr size_t kKeptSearches = 4;
static std::list<OneSeriesBatch> g_kept;

// Frees the device memory of a kept batch (on its own device).
static void drop_one_series_batch(OneSeriesBatch& c)
{
    int cur = 0;
    (void)hipGetDevice(&cur);
    const bool other = c.device >= 0 && c.device != cur;
    if (other) (void)hipSetDevice(c.device);
    if (c.handle) rt_batch_destroy(c.handle);
    if (c.d_periods) (void)hipFree(c.d_periods);
    if (c.d_foldbins) (void)hipFree(c.d_foldbins);
    if (other) (void)hipSetDevice(cur);
    c = OneSeriesBatch();
}

// Drops the kept batches of device `dev` (all devices: -1).
static void drop_kept_batches(int dev)
{
    for (auto it = g_kept.begin(); it != g_kept.end();) {
        if (dev < 0 || it->device == dev) {
            drop_one_series_batch(*it);
            it = g_kept.erase(it);
        } else {
            ++it;
        }
    }
}

// Makes the batch and grids of a search into c: RT_OK, RT_EINVAL when the
// batch path does not take the search, RT_EDEVICE when device memory runs
// out (c left empty but for those).
static int make_one_series_batch(OneSeriesBatch& c, const SeriesPlan& plan, int64_t size,
                                  double tsamp, const uint64_t* h_widths, int64_t nw,
                                  double pmin, double pmax, int64_t bmin, int64_t bmax)
{
    void* h = nullptr;
    const int64_t count = nw;
    const int st = rt_batch_create(1, &size, &tsamp, h_widths, &count, pmin, pmax, &bmin, &bmax,
                                   0, 0.0, 0, 0, 1, &h);
    if (st != RT_OK) {
        (void)hipGetLastError();  // a failed allocation must not fail later launches
        return st == RT_EINVAL ? RT_EINVAL : RT_EDEVICE;
    }
    (void)hipGetDevice(&c.device);
    c.handle = h;
    std::vector<double> periods((size_t)plan.length);
    std::vector<uint32_t> foldbins((size_t)plan.length);
    plan_periods(plan, periods.data(), foldbins.data());
    if (hipMalloc(&c.d_periods, periods.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&c.d_foldbins, foldbins.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMemcpy(c.d_periods, periods.data(), periods.size() * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.d_foldbins, foldbins.data(), foldbins.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
        drop_one_series_batch(c);
        (void)hipGetLastError();
        return RT_EDEVICE;
    }
    c.length = plan.length;
    return RT_OK;
}

// The batch of this search in *out (made on a miss), or null when the batch
// path does not take the search (then the caller runs it pass by pass).
// Returns an error code for a search that is not valid, as the pass path.
static int one_series_batch(int64_t size, double tsamp, const uint64_t* h_widths, int64_t nw,
                            double pmin, double pmax, int64_t bmin, int64_t bmax,
                            OneSeriesBatch** out)
{
    *out = nullptr;
    int device = 0;
    RT_HIP(hipGetDevice(&device));
    const int shape = current_ffa_shape();
    std::vector<uint64_t> w(h_widths, h_widths + std::max<int64_t>(nw, 0));
    std::vector<int32_t> tiling = batch_tiling_flat();
    for (auto it = g_kept.begin(); it != g_kept.end(); ++it) {
        const OneSeriesBatch& c = *it;
        if (c.device == device && c.size == size && c.tsamp == tsamp && c.pmin == pmin &&
            c.pmax == pmax && c.bmin == bmin && c.bmax == bmax && c.shape == shape &&
            c.widths == w && c.tiling == tiling) {
            g_kept.splice(g_kept.begin(), g_kept, it);
            *out = &g_kept.front();
            return RT_OK;
        }
    }
    SeriesPlan plan;
    const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &plan);
    if (!e.empty()) return fail(RT_EINVAL, e);
    const int st = check_widths(h_widths, nw, bmin);
    if (st) return st;
    if (plan.length == 0 || nw == 0) return RT_OK;
    if (g_kept.size() >= kKeptSearches) {
        drop_one_series_batch(g_kept.back());
        g_kept.pop_back();
    }
    OneSeriesBatch c;
    int made = make_one_series_batch(c, plan, size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax);
    if (made == RT_EDEVICE && !g_kept.empty()) {
        // out of device memory: without the kept batches of this device, once more
        drop_kept_batches(device);
        made = make_one_series_batch(c, plan, size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax);
    }
    if (made != RT_OK) return RT_OK;  // pass by pass
    c.size = size;
    c.tsamp = tsamp;
    c.pmin = pmin;
    c.pmax = pmax;
    c.bmin = bmin;
    c.bmax = bmax;
    c.shape = shape;
    c.widths = std::move(w);
    c.tiling = std::move(tiling);
    g_kept.push_front(std::move(c));
    *out = &g_kept.front();
    return RT_OK;
}

}  // namespace rt

extern "C" {

int rt_periodogram(const float* d_in, int64_t size, double tsamp, const uint64_t* h_widths,
                   int64_t nw, double pmin, double pmax, int64_t bmin, int64_t bmax,
                   double* d_periods, uint32_t* d_foldbins, float* d_snrs, void* stream)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    {
        // the batch of one series, where the batch path takes the search
        OneSeriesBatch* ob = nullptr;
        const int st = one_series_batch(size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax, &ob);
        if (st) return st;
        if (ob) {
            const int64_t L = ob->length;
            RT_HIP(hipMemcpyAsync(d_periods, ob->d_periods, L * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
            RT_HIP(hipMemcpyAsync(d_foldbins, ob->d_foldbins, L * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, s));
            const float* in = d_in;
            float* out = d_snrs;
            int bst = rt_batch_bind(ob->handle, &in, &out);
            if (!bst) bst = rt_batch_run(ob->handle, 0, s);
            const int cst = rt_batch_check(ob->handle, s);
            return bst ? bst : cst;
        }
    }
    PgramSetup su;
    int st = pgram_setup(d_in, size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax, d_periods,
                         d_foldbins, s, &su);
    if (st) return st;
    if (su.plan.length == 0 || nw == 0) {
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    // every step at once: pass k of all transforms in one launch
    std::vector<std::vector<FfaTask>> passes;
    const int ns = (int)su.plan.steps.size();
    const int64_t need = pgram_tasks(su, d_in, 0, ns, d_snrs, nullptr, passes, true);
    void* inter = nullptr;
    st = pool_get(kSlotInter, (size_t)std::max<int64_t>(need, 1) * sizeof(float), &inter);
    if (st) return st;
    if (pgram_tasks(su, d_in, 0, ns, d_snrs, (float*)inter, passes, false) < 0)
        return fail(RT_EINVAL, "bins too large for the FFA tiling");
    return run_passes(passes, s, true, su.shape);
}

int rt_release_scratch(void)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    int dev = 0;
    RT_HIP(hipGetDevice(&dev));
    RT_HIP(hipDeviceSynchronize());
    drop_kept_batches(dev);
    auto it = g_pool.find(dev);
    if (it != g_pool.end()) {
        hipError_t err = hipSuccess;
        for (PoolBuf& b : it->second) {
            if (b.p) {
                const hipError_t e = hipFree(b.p);
                if (e != hipSuccess) err = e;
            }
            b = PoolBuf();
        }
        g_pool.erase(it);
        RT_HIP(err);
    }
    return RT_OK;
}

int rt_periodogram_plain(const float* d_in, int64_t size, double tsamp, const uint64_t* h_widths,
                         int64_t nw, double pmin, double pmax, int64_t bmin, int64_t bmax,
                         double* d_periods, uint32_t* d_foldbins, float* d_snrs, void* stream)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    PgramSetup su;
    int st = pgram_setup(d_in, size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax, d_periods,
                         d_foldbins, s, &su);
    if (st) return st;
    if (su.plan.length == 0 || nw == 0) {
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    const int ns = (int)su.plan.steps.size();
    int64_t need = 0;
    for (int is = 0; is < ns; ++is) {
        std::vector<std::vector<FfaTask>> dummy;
        need = std::max(need, pgram_tasks(su, d_in, is, is + 1, d_snrs, nullptr, dummy, true));
    }
    void* inter = nullptr;
    st = pool_get(kSlotInter, (size_t)std::max<int64_t>(need, 1) * sizeof(float), &inter);
    if (st) return st;
    // one launch sequence per step: tasks of all steps are uploaded at once,
    // each step's passes launched in order, steps in order
    std::vector<std::vector<std::vector<FfaTask>>> per_step(ns);
    std::vector<FfaTask> all;
    for (int is = 0; is < ns; ++is)
        if (pgram_tasks(su, d_in, is, is + 1, d_snrs, (float*)inter, per_step[is], false) < 0)
            return fail(RT_EINVAL, "bins too large for the FFA tiling");
    std::vector<std::vector<FfaTask>> flat;
    for (auto& ps : per_step)
        for (auto& p : ps) flat.push_back(p);
    return run_passes(flat, s, true, su.shape);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free periodogram
//
// `scratch` is the only device memory rt_periodogram_noalloc uses besides the
// caller's arrays: the S/N tables, the downsampling descriptors, the packed
// downsampled series, the FFA intermediate states, the task list and the
// error flag are carved from it on every call.  The array's name is a
// host-side stand-in for it: kernels and copies get the device address that
// hipGetSymbolAddress returns, looked up once per device.  Every device the
// library runs on holds its own copy until the process ends.
// ---------------------------------------------------------------------------

namespace rt {

constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
// part of it kept for the task list (about 240k tasks)
constexpr size_t kNoAllocTaskBytes = (size_t)24 << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::map<int, char*> g_scratch_dev;  // device -> address of `scratch`

static int scratch_address(char** out)
{
    int dev = 0;
    RT_HIP(hipGetDevice(&dev));
    auto it = g_scratch_dev.find(dev);
    if (it == g_scratch_dev.end()) {
        void* p = nullptr;
        RT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail(RT_EDEVICE, "no device address for the scratch array");
        it = g_scratch_dev.emplace(dev, (char*)p).first;
    }
    *out = it->second;
    return RT_OK;
}

}  // namespace rt

extern "C" {

int rt_periodogram_noalloc(const float* d_in, int64_t size, double tsamp,
                           const uint64_t* h_widths, int64_t nw, double pmin, double pmax,
                           int64_t bmin, int64_t bmax, double* d_periods, uint32_t* d_foldbins,
                           float* d_snrs, void* stream)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    // sizes first, on the host: nothing runs unless everything fits
    {
        SeriesPlan p;
        const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &p);
        if (!e.empty()) return fail(RT_EINVAL, e);
        int st = check_widths(h_widths, nw, bmin);
        if (st) return st;
        const int shape = current_ffa_shape();
        int64_t big = 0;  // largest intermediate need of one transform (floats)
        for (const Transform& t : p.tr)
            if (t.rows_eval > 0 && t.rows > ffa_cap_rows(t.bins, shape))
                big = std::max<int64_t>(big, 2 * (int64_t)t.rows * t.bins);
        const size_t fixed = 4096 + (size_t)((bmax - bmin + 1) * nw * 2 + nw) * 4 +
                             sizeof(DsSeriesDev) + p.steps.size() * sizeof(DsStepDev) +
                             (size_t)std::max<int64_t>(p.ds_total, 1) * 4 + kNoAllocTaskBytes;
        if (fixed + (size_t)big * 4 > kNoAllocScratchBytes)
            return fail(RT_EINVAL,
                        "periodogram needs " + std::to_string((fixed + big * 4) >> 20) +
                            " MiB of scratch, more than the " +
                            std::to_string(kNoAllocScratchBytes >> 20) +
                            " MiB the allocation-free entry point holds");
    }
    char* base = nullptr;
    int st = scratch_address(&base);
    if (st) return st;
    // [tables | downsampling descriptors | downsampled series | intermediates][task list]
    Arena a{base, kNoAllocScratchBytes - kNoAllocTaskBytes, 0};
    Arena ta{base + (kNoAllocScratchBytes - kNoAllocTaskBytes), kNoAllocTaskBytes, 0};
    PgramSetup su;
    st = pgram_setup(d_in, size, tsamp, h_widths, nw, pmin, pmax, bmin, bmax, d_periods,
                     d_foldbins, s, &su, &a);
    if (st) return st;
    if (su.plan.length == 0 || nw == 0) {
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    // the rest holds the intermediate states of a chunk of transforms; the
    // chunks run one after the other on the stream, each reusing it
    const size_t off = (a.used + 255) & ~(size_t)255;
    const int64_t inter_floats = (int64_t)((a.cap - std::min(a.cap, off)) / 4);
    void* ip = nullptr;
    st = arena_get(&a, (size_t)inter_floats * 4, &ip);
    if (st) return st;
    float* inter = (float*)ip;
    std::vector<std::vector<FfaTask>> all, chunk;
    int64_t used = 0;
    auto flush = [&]() {
        for (auto& q : chunk) all.push_back(std::move(q));
        chunk.clear();
        used = 0;
    };
    const SeriesPlan& p = su.plan;
    for (const Step& stp : p.steps)
        for (int it = stp.first_tr; it < stp.first_tr + stp.ntr; ++it) {
            const Transform& t = p.tr[it];
            if (t.rows_eval <= 0) continue;
            TransformSpec sp = pgram_spec(su, d_in, stp, t, d_snrs);
            if (t.rows > sp.cap) {
                const int64_t n = (int64_t)t.rows * t.bins;
                if (2 * n > inter_floats)
                    return fail(RT_EINVAL, "FFA intermediate states exceed the scratch array");
                if (used + 2 * n > inter_floats) flush();
                sp.buf_a = inter + used;
                sp.buf_b = inter + used + n;
                used += 2 * n;
            }
            if (plan_transform_tasks(sp, chunk) < 0)
                return fail(RT_EINVAL, "bins too large for the FFA tiling");
        }
    flush();
    return run_passes(all, s, true, su.shape, &ta);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoints
// ---------------------------------------------------------------------------

extern "C" {

int rt_normalise_ex(const float* d_in, int64_t size, float* d_out, float* h_stats, void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "cannot normalise an empty series");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    void* p = nullptr;
    int st = pool_get(kSlotMisc, (size_t)(2 * (kNormaliseParts + 2)) * sizeof(float), &p);
    if (st) return st;
    RT_HIP(launch_normalise(d_in, size, d_out, (float*)p, kNormaliseParts, s));
    // mean (hi, lo) and std (hi, lo) sit after the partials
    RT_HIP(hipMemcpyAsync(h_stats, (float*)p + 2 * kNormaliseParts, 4 * sizeof(float),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_normalise_backward(const float* d_g, const float* d_y, int64_t size, float inv_std,
                          float* d_gx, void* stream)
{
    if (size < 1) return fail(RT_EINVAL, "empty series");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    void* p = nullptr;
    const int parts = 256;
    int st = pool_get(kSlotMisc, (size_t)(4 * parts) * sizeof(float), &p);
    if (st) return st;
    RT_HIP(launch_normalise_adjoint(d_g, d_y, size, inv_std, (float*)p, parts, d_gx, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_ffa2_backward(const float* d_gout, int64_t rows, int64_t cols, float* d_gin, void* stream)
{
    if (rows < 1 || cols < 1)
        return fail(RT_EINVAL, "ffa2 input must have at least one row and one column");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    const int H = ffa_height(rows);
    const size_t n = (size_t)(rows * cols);
    if (H == 0) {
        RT_HIP(hipMemcpyAsync(d_gin, d_gout, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    } else {
        void* tmp = nullptr;
        int st = pool_get(kSlotInter, n * sizeof(float), &tmp);
        if (st) return st;
        const float* src = d_gout;
        for (int d = 0; d < H; ++d) {
            const bool to_out = ((H - 1 - d) % 2) == 0;
            float* dst = to_out ? d_gin : (float*)tmp;
            RT_HIP(launch_ffa_adjoint_level(src, dst, rows, cols, d, s));
            src = dst;
        }
    }
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_snr2_backward(const float* d_in, int64_t rows, int64_t cols, const uint64_t* h_widths,
                     int64_t nw, float stdnoise, const float* d_gsnr, float* d_gin, void* stream)
{
    int st = check_stdnoise(stdnoise);
    if (st) return st;
    st = check_widths(h_widths, nw, cols);
    if (st) return st;
    if (nw > kSnrBwdMaxW) return fail(RT_EINVAL, "at most 64 trial widths in the S/N backward");
    if (rows == 0) return RT_OK;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    std::vector<float> tab = hb_table(cols, 1, h_widths, nw);
    const size_t woff = tab.size();
    tab.resize(woff + (size_t)nw);
    for (int64_t i = 0; i < nw; ++i) {
        const int32_t w = (int32_t)h_widths[i];
        std::memcpy(&tab[woff + (size_t)i], &w, sizeof(int32_t));
    }
    float* d_tab = nullptr;
    st = upload(kSlotTables, tab, &d_tab, s);
    if (st) return st;
    RT_HIP(launch_snr_backward(d_in, (int)rows, (int)cols, d_gsnr, (const int32_t*)(d_tab + woff),
                               d_tab, (int)nw, stdnoise, d_gin, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

int rt_periodogram_backward(const float* d_in, int64_t size, double tsamp,
                            const uint64_t* h_widths, int64_t nw, double pmin, double pmax,
                            int64_t bmin, int64_t bmax, const float* d_gsnr, float* d_gin,
                            void* stream)
{
    if (nw > kSnrBwdMaxW) return fail(RT_EINVAL, "at most 64 trial widths in the S/N backward");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    hipStream_t s = S(stream);
    SeriesPlan p;
    const std::string e = plan_series(size, tsamp, pmin, pmax, bmin, bmax, &p);
    if (!e.empty()) return fail(RT_EINVAL, e);
    int st = check_widths(h_widths, nw, bmin);
    if (st) return st;
    RT_HIP(launch_fill_zero(d_gin, size, s));
    if (p.length == 0 || nw == 0) {
        RT_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    // which transforms receive any gradient: flags computed on the device
    // from each transform's rows of d_gsnr (only the flags cross to the host)
    const int ntr = (int)p.tr.size();
    std::vector<int32_t> meta((size_t)3 * ntr);
    for (int it = 0; it < ntr; ++it) {
        meta[2 * it] = (int32_t)p.tr[it].out_row;
        meta[2 * it + 1] = (int32_t)p.tr[it].rows_eval;
    }
    int32_t* d_meta = nullptr;
    st = upload(kSlotMisc, meta, &d_meta, s);
    if (st) return st;
    int32_t* d_flags = d_meta + 2 * ntr;
    RT_HIP(launch_transform_need(d_gsnr, reinterpret_cast<const int2*>(d_meta), ntr, (int)nw,
                                 d_flags, s));
    std::vector<int32_t> flags(ntr);
    RT_HIP(hipMemcpyAsync(flags.data(), d_flags, (size_t)ntr * sizeof(int32_t),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    std::vector<char> need_tr(p.tr.size(), 0), need_step(p.steps.size(), 0);
    for (int it = 0; it < ntr; ++it)
        if (flags[it]) {
            need_tr[it] = 1;
            need_step[p.tr[it].step] = 1;
        }
    // S/N tables for bins in [bmin, bmax]
    std::vector<float> tab = hb_table(bmin, bmax - bmin + 1, h_widths, nw);
    const size_t woff = tab.size();
    tab.resize(woff + (size_t)nw);
    for (int64_t i = 0; i < nw; ++i) {
        const int32_t w = (int32_t)h_widths[i];
        std::memcpy(&tab[woff + (size_t)i], &w, sizeof(int32_t));
    }
    float* d_tab = nullptr;
    st = upload(kSlotTables, tab, &d_tab, s);
    if (st) return st;
    // workspace: ds values + ds gradient + downsampling index table (max n)
    // and three transform buffers
    int64_t max_fold = 1;
    for (size_t it = 0; it < p.tr.size(); ++it)
        if (need_tr[it]) max_fold = std::max<int64_t>(max_fold, (int64_t)p.tr[it].rows * p.tr[it].bins);
    void* wsp = nullptr;
    st = pool_get(kSlotInter, (size_t)(6 * p.max_n + 4 * max_fold) * sizeof(float), &wsp);
    if (st) return st;
    // the index table first: int4 entries need 16-byte alignment
    int4* dsidx = reinterpret_cast<int4*>(wsp);
    float* ds = reinterpret_cast<float*>(dsidx + p.max_n);
    float* gds = ds + p.max_n;
    float* T = gds + p.max_n;
    float* A = T + max_fold;
    float* B = A + max_fold;
    float* C = B + max_fold;
    void* errp = nullptr;
    st = pool_get(kSlotErr, sizeof(int), &errp);
    if (st) return st;
    RT_HIP(hipMemsetAsync(errp, 0, sizeof(int), s));

    // the forward tasks of every transform to recompute, uploaded at once
    const int shape = current_ffa_shape();
    std::vector<FfaTask> all;
    std::vector<std::vector<size_t>> pass_ends(p.tr.size());
    std::vector<size_t> tr_first(p.tr.size(), 0);
    for (size_t is = 0; is < p.steps.size(); ++is) {
        if (!need_step[is]) continue;
        const Step& stp = p.steps[is];
        const float* x = stp.ds_offset >= 0 ? ds : d_in;
        for (int it = stp.first_tr; it < stp.first_tr + stp.ntr; ++it) {
            if (!need_tr[it]) continue;
            const Transform& t = p.tr[it];
            TransformSpec sp;
            sp.m = t.rows;
            sp.bins = t.bins;
            sp.cap = ffa_cap_rows(t.bins, shape);
            sp.shape = shape;
            sp.src = x;
            sp.kind = 0;
            sp.dst = T;
            sp.buf_a = A;
            sp.buf_b = B;
            std::vector<std::vector<FfaTask>> passes;
            if (plan_transform_tasks(sp, passes) < 0)
                return fail(RT_EINVAL, "bins too large for the FFA tiling");
            tr_first[it] = all.size();
            for (auto& q : passes) {
                all.insert(all.end(), q.begin(), q.end());
                pass_ends[it].push_back(all.size());
            }
        }
    }
    FfaTask* d_tasks = nullptr;
    st = upload(kSlotTasks, all, &d_tasks, s);
    if (st) return st;

    for (size_t is = 0; is < p.steps.size(); ++is) {
        if (!need_step[is]) continue;
        const Step& stp = p.steps[is];
        if (stp.ds_offset >= 0) RT_HIP(launch_downsample_direct(d_in, size, stp.f, stp.n, ds, s));
        RT_HIP(launch_fill_zero(gds, stp.n, s));
        for (int it = stp.first_tr; it < stp.first_tr + stp.ntr; ++it) {
            if (!need_tr[it]) continue;
            const Transform& t = p.tr[it];
            const int64_t n = (int64_t)t.rows * t.bins;
            // forward transform rows (recomputed)
            size_t off = tr_first[it];
            for (size_t e : pass_ends[it]) {
                RT_HIP(launch_ffa_task_list(d_tasks + off, all.data() + off, (int)(e - off),
                                            (int*)errp, s, shape));
                off = e;
            }
            // dL/d(transform rows): S/N backward on the evaluated rows, zero beyond
            RT_HIP(launch_fill_zero(A, n, s));
            RT_HIP(launch_snr_backward(T, t.rows_eval, t.bins, d_gsnr + t.out_row * nw,
                                       (const int32_t*)(d_tab + woff),
                                       d_tab + (size_t)(t.bins - bmin) * nw * 2, (int)nw,
                                       t.stdnoise, A, s));
            // FFA transpose, depth 0 .. H-1
            const int H = ffa_height(t.rows);
            const float* src = A;
            float* bufs[2] = {B, C};
            for (int d = 0; d < H; ++d) {
                float* dst = bufs[d & 1];
                RT_HIP(launch_ffa_adjoint_level(src, dst, t.rows, t.bins, d, s));
                src = dst;
            }
            // fold transpose: the first rows*bins samples, accumulated
            RT_HIP(launch_add_prefix(src, n, stp.n, gds, 1, s));
        }
        if (stp.ds_offset >= 0)
            RT_HIP(launch_downsample_adjoint(gds, stp.n, stp.f, size, d_gin, 1, dsidx, s));
        else
            RT_HIP(launch_add_prefix(gds, stp.n, size, d_gin, 1, s));
    }
    int herr = 0;
    RT_HIP(hipMemcpyAsync(&herr, errp, sizeof(int), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (herr & 1) return fail(RT_EDEVICE, "FFA tile exceeded its LDS budget");
    if (herr) return fail(RT_EDEVICE, "FFA task launched with the wrong kernel variant");
    return RT_OK;
}

}  // extern "C"
