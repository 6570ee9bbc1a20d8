// Batched differentiable periodogram: normalisation and backward of many
// series of one length that share one search plan (the forward periodogram
// itself runs as a batch, rt_batch_*).  The backward takes a fixed number of
// launches for the whole batch, whatever the number of series: one device
// pass finds the (series, transform) pairs that receive gradient, then every
// stage (downsampling index tables and values, forward recompute of the
// pairs' transforms, S/N backward, each transposed merge level, fold and
// downsampling transposes, normalisation transpose) is one launch over all
// pairs or groups.  Per element, the arithmetic and its order are those of
// rt_periodogram_backward + rt_normalise_backward on one series, so the
// gradients are bit-identical to the one-series path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/riptide_amd.h"
#include "adjoint.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "plan.h"
#include "series_kernels.h"
#include "tasks.h"

namespace rt {

void set_error(const std::string& msg);
int current_ffa_shape();

namespace {

int gfail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

#define GB_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return gfail(RT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Blocks per item of the batched element-wise kernels (each block strides
// over the item's elements); a few per item keep the block tables small.
constexpr int kItemBlocks = 64;
// Partials of the normalisation transpose (the one-series path's count).
constexpr int kNormAdjParts = 256;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int blocks_for(int64_t n, int threads, int cap)
{
    const int64_t b = (n + threads - 1) / threads;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

}  // namespace

struct GradBatch {
    int device = 0;
    int64_t nseries = 0, size = 0;
    double tsamp = 0, pmin = 0, pmax = 0;
    int64_t bmin = 0, bmax = 0, nw = 0;
    SeriesPlan plan;
    int shape = 0;
    // fixed device buffers
    float* d_tab = nullptr;          // (h + b, b) per (bins, width), then the widths (int32)
    int64_t woff = 0;
    int2* d_rows = nullptr;          // (out_row, rows_eval) per transform
    int* d_flags = nullptr;          // nseries x transforms
    int* d_err = nullptr;
    float* d_norm = nullptr;         // per series: normalisation partials, mean, std
    PrepSeriesDev* d_prep = nullptr;
    float* d_inv_std = nullptr;      // per series, of the last normalisation
    float* d_part = nullptr;         // normalisation-transpose partials per series
    float* d_gz = nullptr;           // gradient w.r.t. the normalised series
    // grow-only work areas
    char* d_work = nullptr;
    size_t work_cap = 0;
    char* d_meta = nullptr;
    size_t meta_cap = 0;
    char* h_meta = nullptr;          // pinned staging of the tables
    size_t h_meta_cap = 0;
    std::vector<float> inv_std;
    std::vector<PrepSeriesDev> prep;
    std::vector<int> flags;
};

static void free_grad(GradBatch* g)
{
    for (void* p : {(void*)g->d_tab, (void*)g->d_rows, (void*)g->d_flags, (void*)g->d_err,
                    (void*)g->d_norm, (void*)g->d_prep, (void*)g->d_inv_std, (void*)g->d_part,
                    (void*)g->d_gz, (void*)g->d_work, (void*)g->d_meta})
        if (p) (void)hipFree(p);
    if (g->h_meta) (void)hipHostFree(g->h_meta);
    delete g;
}

static int grow_device(char** p, size_t* cap, size_t need)
{
    if (need <= *cap) return RT_OK;
    if (*p) {
        GB_HIP(hipDeviceSynchronize());
        GB_HIP(hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    GB_HIP(hipMalloc(p, need));
    *cap = need;
    return RT_OK;
}

static int grow_host(char** p, size_t* cap, size_t need)
{
    if (need <= *cap) return RT_OK;
    if (*p) {
        GB_HIP(hipDeviceSynchronize());
        GB_HIP(hipHostFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    GB_HIP(hipHostMalloc(p, need, hipHostMallocDefault));
    *cap = need;
    return RT_OK;
}

// Tables of one backward, packed into one host buffer and copied at once.
struct MetaPack {
    std::vector<char> bytes;
    template <typename T>
    size_t add(const std::vector<T>& v)
    {
        const size_t off = align256(bytes.size());
        bytes.resize(off + std::max<size_t>(v.size(), 1) * sizeof(T));
        if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

}  // namespace rt

using namespace rt;

extern "C" {

int rt_grad_batch_create(int64_t nseries, int64_t size, double tsamp, const uint64_t* h_widths,
                         int64_t nw, double period_min, double period_max, int64_t bins_min,
                         int64_t bins_max, void** handle)
{
    if (nseries < 1) return gfail(RT_EINVAL, "batch needs at least one series");
    if (nw < 1 || nw > kSnrBwdMaxW) return gfail(RT_EINVAL, "1 to 64 trial widths");
    for (int64_t i = 0; i < nw; ++i)
        if (!(h_widths[i] > 0 && h_widths[i] < (uint64_t)bins_min))
            return gfail(RT_EINVAL, "trial widths must be all > 0 and < columns");
    GradBatch* g = new GradBatch();
    const std::string e = plan_series(size, tsamp, period_min, period_max, bins_min, bins_max,
                                      &g->plan);
    if (!e.empty()) {
        free_grad(g);
        return gfail(RT_EINVAL, e);
    }
    GB_HIP(hipGetDevice(&g->device));
    g->nseries = nseries;
    g->size = size;
    g->tsamp = tsamp;
    g->pmin = period_min;
    g->pmax = period_max;
    g->bmin = bins_min;
    g->bmax = bins_max;
    g->nw = nw;
    g->shape = current_ffa_shape();
    // S/N tables
    std::vector<float> tab;
    for (int64_t b = bins_min; b <= bins_max; ++b)
        for (int64_t i = 0; i < nw; ++i) {
            float hpb = 0, bb = 0;
            if (h_widths[i] < (uint64_t)b) boxcar_hb(b, (int64_t)h_widths[i], &hpb, &bb);
            tab.push_back(hpb);
            tab.push_back(bb);
        }
    g->woff = (int64_t)tab.size();
    for (int64_t i = 0; i < nw; ++i) {
        const int32_t w = (int32_t)h_widths[i];
        float f;
        std::memcpy(&f, &w, sizeof(f));
        tab.push_back(f);
    }
    const int ntr = (int)g->plan.tr.size();
    std::vector<int2> rows(std::max(ntr, 1));
    for (int it = 0; it < ntr; ++it)
        rows[it] = make_int2((int)g->plan.tr[it].out_row, (int)g->plan.tr[it].rows_eval);
    const size_t nstat = 2 * kNormaliseParts + 4;
    GB_HIP(hipMalloc(&g->d_tab, tab.size() * sizeof(float)));
    GB_HIP(hipMemcpy(g->d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    GB_HIP(hipMalloc(&g->d_rows, rows.size() * sizeof(int2)));
    GB_HIP(hipMemcpy(g->d_rows, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
    GB_HIP(hipMalloc(&g->d_flags, (size_t)nseries * std::max(ntr, 1) * sizeof(int)));
    GB_HIP(hipMalloc(&g->d_err, sizeof(int)));
    GB_HIP(hipMalloc(&g->d_norm, (size_t)nseries * nstat * sizeof(float)));
    GB_HIP(hipMalloc(&g->d_prep, (size_t)nseries * sizeof(PrepSeriesDev)));
    GB_HIP(hipMalloc(&g->d_inv_std, (size_t)nseries * sizeof(float)));
    GB_HIP(hipMalloc(&g->d_part, (size_t)nseries * 4 * kNormAdjParts * sizeof(float)));
    GB_HIP(hipMalloc(&g->d_gz, (size_t)nseries * size * sizeof(float)));
    g->inv_std.assign(nseries, 1.0f);
    g->prep.assign(nseries, PrepSeriesDev{});
    *handle = g;
    return RT_OK;
}

int rt_grad_batch_length(void* handle, int64_t* length)
{
    *length = ((GradBatch*)handle)->plan.length;
    return RT_OK;
}

int rt_grad_batch_normalise(void* handle, const float* d_x, float* d_y, float* h_inv_std,
                            void* stream)
{
    GradBatch* g = (GradBatch*)handle;
    hipStream_t s = (hipStream_t)stream;
    const size_t nstat = 2 * kNormaliseParts + 4;
    for (int64_t i = 0; i < g->nseries; ++i) {
        PrepSeriesDev& d = g->prep[i];
        d = PrepSeriesDev{};
        d.x = d_x + i * g->size;
        d.out = d_y + i * g->size;
        d.stats = g->d_norm + i * nstat;
        d.N = g->size;
        d.nl = g->size;
        d.sf = 1;
        d.width = 1;
        d.dereddened = 0;
    }
    GB_HIP(hipMemcpyAsync(g->d_prep, g->prep.data(), g->prep.size() * sizeof(PrepSeriesDev),
                          hipMemcpyHostToDevice, s));
    GB_HIP(launch_prep_batch(g->d_prep, (int)g->nseries, g->size, 0, 1, 0, 1, s));
    // 1/std per series, as the one-series path derives it on the host
    std::vector<float> stats((size_t)g->nseries * nstat);
    GB_HIP(hipMemcpyAsync(stats.data(), g->d_norm, stats.size() * sizeof(float),
                          hipMemcpyDeviceToHost, s));
    GB_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < g->nseries; ++i) {
        const float* st = stats.data() + i * nstat + 2 * kNormaliseParts;
        g->inv_std[i] = (float)(1.0 / ((double)st[2] + (double)st[3]));
        if (h_inv_std) h_inv_std[i] = g->inv_std[i];
    }
    GB_HIP(hipMemcpy(g->d_inv_std, g->inv_std.data(), g->inv_std.size() * sizeof(float),
                     hipMemcpyHostToDevice));
    return RT_OK;
}

int rt_grad_batch_backward(void* handle, const float* d_y, const float* h_inv_std,
                           const float* d_gsnr, float* d_gx, void* stream)
{
    GradBatch* g = (GradBatch*)handle;
    hipStream_t s = (hipStream_t)stream;
    std::copy(h_inv_std, h_inv_std + g->nseries, g->inv_std.begin());
    GB_HIP(hipMemcpyAsync(g->d_inv_std, g->inv_std.data(), g->inv_std.size() * sizeof(float),
                          hipMemcpyHostToDevice, s));
    const SeriesPlan& p = g->plan;
    const int ntr = (int)p.tr.size();
    const int64_t N = g->size;
    const int nw = (int)g->nw;
    if (p.length == 0 || ntr == 0) {
        GB_HIP(hipMemsetAsync(d_gx, 0, (size_t)g->nseries * N * sizeof(float), s));
        GB_HIP(hipStreamSynchronize(s));
        return RT_OK;
    }
    // (1) the (series, transform) pairs that receive gradient
    GB_HIP(launch_flags_batch(d_gsnr, p.length * g->nw, g->d_rows, ntr, nw, (int)g->nseries,
                              g->d_flags, s));
    g->flags.resize((size_t)g->nseries * ntr);
    GB_HIP(hipMemcpyAsync(g->flags.data(), g->d_flags, g->flags.size() * sizeof(int),
                          hipMemcpyDeviceToHost, s));
    GB_HIP(hipMemsetAsync(g->d_err, 0, sizeof(int), s));
    GB_HIP(hipStreamSynchronize(s));

    // (2) pairs, (series, step) groups and the steps they use
    struct PairH { int series, tr, group; };
    struct GroupH { int series, step; int p0, p1; };
    std::vector<PairH> ph;
    std::vector<GroupH> gh;
    std::vector<int> g0((size_t)g->nseries + 1, 0);
    std::vector<int> step_slot(p.steps.size(), -1);
    std::vector<int> used_steps;
    for (int64_t si = 0; si < g->nseries; ++si) {
        g0[si] = (int)gh.size();
        for (size_t is = 0; is < p.steps.size(); ++is) {
            const Step& st = p.steps[is];
            const int first = (int)ph.size();
            for (int it = st.first_tr; it < st.first_tr + st.ntr; ++it)
                if (g->flags[(size_t)si * ntr + it] && p.tr[it].rows_eval > 0)
                    ph.push_back({(int)si, it, (int)gh.size()});
            if ((int)ph.size() > first) {
                if (step_slot[is] < 0) {
                    step_slot[is] = (int)used_steps.size();
                    used_steps.push_back((int)is);
                }
                gh.push_back({(int)si, step_slot[is], first, (int)ph.size()});
            }
        }
    }
    g0[g->nseries] = (int)gh.size();

    // (3) device work area: index tables, downsampled values and fold
    // gradients per group, four transform buffers per pair
    size_t off = 0;
    std::vector<size_t> idx_off(used_steps.size()), ds_off(gh.size()), gds_off(gh.size());
    std::vector<size_t> pair_off(ph.size());
    for (size_t k = 0; k < used_steps.size(); ++k) {
        const Step& st = p.steps[used_steps[k]];
        idx_off[k] = off;
        if (st.ds_offset >= 0) off = align256(off + (size_t)st.n * sizeof(int4));
    }
    for (size_t k = 0; k < gh.size(); ++k) {
        const Step& st = p.steps[used_steps[gh[k].step]];
        ds_off[k] = off;
        if (st.ds_offset >= 0) off = align256(off + (size_t)st.n * sizeof(float));
        gds_off[k] = off;
        off = align256(off + (size_t)st.n * sizeof(float));
    }
    for (size_t k = 0; k < ph.size(); ++k) {
        const Transform& t = p.tr[ph[k].tr];
        pair_off[k] = off;
        off = align256(off + 4 * (size_t)t.rows * t.bins * sizeof(float));
    }
    int st = grow_device(&g->d_work, &g->work_cap, std::max<size_t>(off, 256));
    if (st) return st;
    char* W = g->d_work;

    // (4) descriptors with device addresses
    std::vector<GradStep> steps(used_steps.size());
    for (size_t k = 0; k < used_steps.size(); ++k) {
        const Step& sp = p.steps[used_steps[k]];
        GradStep& d = steps[k];
        d.n = sp.n;
        if (sp.ds_offset >= 0) {
            int ef = 0;
            split_double(sp.f, &d.mf, &ef);
            d.ef = ef;
            d.inv_f = (float)(1.0 / sp.f);
            d.idx = reinterpret_cast<int4*>(W + idx_off[k]);
        } else {
            d.mf = 0;
            d.ef = 0;
            d.inv_f = 1.0f;
            d.idx = nullptr;
        }
    }
    std::vector<GradGroup> groups(gh.size());
    for (size_t k = 0; k < gh.size(); ++k) {
        const Step& sp = p.steps[used_steps[gh[k].step]];
        GradGroup& d = groups[k];
        d.x = d_y + (size_t)gh[k].series * N;
        d.ds = sp.ds_offset >= 0 ? reinterpret_cast<float*>(W + ds_off[k]) : const_cast<float*>(d.x);
        d.gds = reinterpret_cast<float*>(W + gds_off[k]);
        d.step = gh[k].step;
        d.p0 = gh[k].p0;
        d.p1 = gh[k].p1;
        d.pad = 0;
    }
    std::vector<GradPair> pairs(ph.size());
    int max_bins = 1, max_h = 0;
    for (size_t k = 0; k < ph.size(); ++k) {
        const Transform& t = p.tr[ph[k].tr];
        const size_t n = (size_t)t.rows * t.bins;
        float* base = reinterpret_cast<float*>(W + pair_off[k]);
        GradPair& d = pairs[k];
        d.T = base;
        d.A = base + n;
        d.B = base + 2 * n;
        d.C = base + 3 * n;
        d.gsnr = d_gsnr + ((size_t)ph[k].series * p.length + t.out_row) * nw;
        d.hb = g->d_tab + (size_t)(t.bins - g->bmin) * nw * 2;
        d.rows = t.rows;
        d.bins = t.bins;
        d.rows_eval = t.rows_eval;
        d.H = ffa_height(t.rows);
        d.stdnoise = t.stdnoise;
        d.pad = 0;
        max_bins = std::max(max_bins, (int)t.bins);
        max_h = std::max(max_h, (int)d.H);
    }
    // forward recompute tasks: pass q of every pair in list q
    std::vector<std::vector<FfaTask>> passes;
    for (size_t k = 0; k < ph.size(); ++k) {
        const Transform& t = p.tr[ph[k].tr];
        TransformSpec sp;
        sp.m = t.rows;
        sp.bins = t.bins;
        sp.cap = ffa_cap_rows(t.bins, g->shape);
        sp.shape = g->shape;
        sp.src = groups[ph[k].group].ds;
        sp.kind = 0;
        sp.dst = const_cast<float*>(pairs[k].T);
        sp.buf_a = pairs[k].A;
        sp.buf_b = pairs[k].B;
        if (plan_transform_tasks(sp, passes) < 0)
            return gfail(RT_EINVAL, "bins too large for the FFA tiling");
    }
    std::vector<FfaTask> tasks;
    std::vector<size_t> pass_first;
    for (auto& q : passes) {
        pass_first.push_back(tasks.size());
        sort_by_variant(q);
        tasks.insert(tasks.end(), q.begin(), q.end());
    }
    pass_first.push_back(tasks.size());
    // block tables
    std::vector<int4> blk_idx, blk_ds, blk_snr, blk_fold, blk_fill;
    std::vector<std::vector<int4>> blk_lvl((size_t)max_h);
    for (size_t k = 0; k < steps.size(); ++k)
        if (steps[k].idx) {
            const int nb = blocks_for(steps[k].n, 256, kItemBlocks);
            for (int b = 0; b < nb; ++b) blk_idx.push_back(make_int4((int)k, b, nb, 0));
        }
    for (size_t k = 0; k < groups.size(); ++k) {
        const int64_t n = steps[groups[k].step].n;
        const int nb = blocks_for(n, 256, kItemBlocks);
        if (steps[groups[k].step].idx)
            for (int b = 0; b < nb; ++b) blk_ds.push_back(make_int4((int)k, b, nb, 0));
        for (int b = 0; b < nb; ++b) blk_fold.push_back(make_int4((int)k, b, nb, 0));
    }
    for (size_t k = 0; k < pairs.size(); ++k) {
        const int nb = (pairs[k].rows_eval + kGradSnrWaves - 1) / kGradSnrWaves;
        for (int b = 0; b < nb; ++b) blk_snr.push_back(make_int4((int)k, b, nb, 0));
        const int nl = blocks_for((int64_t)pairs[k].rows * pairs[k].bins, 256, kItemBlocks);
        for (int b = 0; b < nl; ++b) blk_fill.push_back(make_int4((int)k, b, nl, 0));
        for (int d = 0; d < pairs[k].H; ++d)
            for (int b = 0; b < nl; ++b) blk_lvl[d].push_back(make_int4((int)k, b, nl, d));
    }
    MetaPack mp;
    const size_t o_tasks = mp.add(tasks), o_steps = mp.add(steps), o_groups = mp.add(groups),
                 o_pairs = mp.add(pairs), o_g0 = mp.add(g0), o_bidx = mp.add(blk_idx),
                 o_bds = mp.add(blk_ds), o_bsnr = mp.add(blk_snr), o_bfold = mp.add(blk_fold),
                 o_bfill = mp.add(blk_fill);
    std::vector<size_t> o_lvl((size_t)max_h);
    for (int d = 0; d < max_h; ++d) o_lvl[d] = mp.add(blk_lvl[d]);
    st = grow_device(&g->d_meta, &g->meta_cap, mp.bytes.size());
    if (st) return st;
    st = grow_host(&g->h_meta, &g->h_meta_cap, mp.bytes.size());
    if (st) return st;
    std::memcpy(g->h_meta, mp.bytes.data(), mp.bytes.size());
    GB_HIP(hipMemcpyAsync(g->d_meta, g->h_meta, mp.bytes.size(), hipMemcpyHostToDevice, s));
    char* M = g->d_meta;
    const FfaTask* d_tasks = reinterpret_cast<const FfaTask*>(M + o_tasks);
    const GradStep* d_steps = reinterpret_cast<const GradStep*>(M + o_steps);
    const GradGroup* d_groups = reinterpret_cast<const GradGroup*>(M + o_groups);
    const GradPair* d_pairs = reinterpret_cast<const GradPair*>(M + o_pairs);
    const int* d_g0 = reinterpret_cast<const int*>(M + o_g0);

    // (5) the launches
    GB_HIP(launch_ds_index_steps(d_steps, reinterpret_cast<const int4*>(M + o_bidx),
                                 (int)blk_idx.size(), N, s));
    GB_HIP(launch_ds_values_groups(d_groups, d_steps, reinterpret_cast<const int4*>(M + o_bds),
                                   (int)blk_ds.size(), s));
    for (size_t q = 0; q + 1 < pass_first.size(); ++q)
        GB_HIP(launch_ffa_task_list(d_tasks + pass_first[q], tasks.data() + pass_first[q],
                                    (int)(pass_first[q + 1] - pass_first[q]), g->d_err, s,
                                    g->shape));
    GB_HIP(launch_fill_pairs(d_pairs, reinterpret_cast<const int4*>(M + o_bfill),
                             (int)blk_fill.size(), s));
    GB_HIP(launch_snr_backward_pairs(d_pairs, reinterpret_cast<const int4*>(M + o_bsnr),
                                     (int)blk_snr.size(), max_bins,
                                     reinterpret_cast<const int32_t*>(g->d_tab + g->woff), nw, s));
    for (int d = 0; d < max_h; ++d)
        GB_HIP(launch_ffa_adjoint_level_pairs(d_pairs, reinterpret_cast<const int4*>(M + o_lvl[d]),
                                              (int)blk_lvl[d].size(), s));
    GB_HIP(launch_fold_groups(d_groups, d_pairs, d_steps, reinterpret_cast<const int4*>(M + o_bfold),
                              (int)blk_fold.size(), s));
    GB_HIP(launch_ds_adjoint_series(d_groups, d_steps, d_g0, (int)g->nseries, N, g->d_gz, s));
    GB_HIP(launch_normalise_adjoint_batch(g->d_gz, d_y, N, (int)g->nseries, g->d_inv_std,
                                          g->d_part, kNormAdjParts, d_gx, s));
    int herr = 0;
    GB_HIP(hipMemcpyAsync(&herr, g->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    GB_HIP(hipStreamSynchronize(s));
    if (herr & 1) return gfail(RT_EDEVICE, "FFA tile exceeded its LDS budget");
    if (herr) return gfail(RT_EDEVICE, "FFA task launched with the wrong kernel variant");
    return RT_OK;
}

int rt_grad_batch_destroy(void* handle)
{
    if (handle) free_grad((GradBatch*)handle);
    return RT_OK;
}

}  // extern "C"
