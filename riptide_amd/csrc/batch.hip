// Batched DM-trial search: many series (uneven lengths, per-series bins
// ranges) searched with a handful of launches per group of series:
//   prepare (dereddening + normalisation) -> one packed downsampling launch
//   for every step of every series -> one persistent FFA launch that drains
//   a device queue of whole transforms, largest first.
// A run only enqueues kernels with fixed arguments, so it can be captured
// once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/riptide_amd.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "plan.h"
#include "series_kernels.h"
#include "task_iter.h"
#include "tasks.h"

namespace rt {

void set_error(const std::string& msg);
int current_ffa_shape();

namespace {

int bfail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

#define BT_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return bfail(RT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__global__ void fill_int_kernel(int* p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Transforms of one group that share a kernel variant: one persistent
// launch, one queue counter.
struct Run {
    int tr0 = 0, ntr = 0;          // transforms in the device table
    int variant = 0;
    int shape = 0;                 // workgroup shape of the launch
    int queue = 0;                 // queue counter index
};

// Tiling of the transforms of one (rows, bins) class, chosen by measurement:
// workgroup shape, rows per tile (0: as many as fit) and merge levels per
// cone pass (0: the planner's limit).
struct TileClass {
    int32_t rows_lo, rows_hi;      // rows in [rows_lo, rows_hi)
    int32_t bins_lo, bins_hi;      // bins in [bins_lo, bins_hi)
    int32_t shape, cap, lmax;
};
std::vector<TileClass> g_tile_classes;

// The tiling of a transform: the first class that holds it, else the
// current shape with the largest tile; rows too wide for that shape's
// staging take the shape with the deepest staging that holds them.
void tile_for(int64_t rows, int64_t bins, int default_shape, int* shape, int* cap, int* lmax)
{
    *shape = default_shape;
    *cap = 0;
    *lmax = kMaxConeLevels;
    for (const TileClass& c : g_tile_classes) {
        if (rows >= c.rows_lo && rows < c.rows_hi && bins >= c.bins_lo && bins < c.bins_hi) {
            *shape = c.shape;
            *cap = c.cap;
            *lmax = c.lmax > 0 ? c.lmax : kMaxConeLevels;
            break;
        }
    }
    int most = ffa_cap_rows((int)bins, *shape);
    if (most < 1) {
        int best = -1;
        for (int sh = 0; sh < kFfaShapes; ++sh)
            if (ffa_cap_rows((int)bins, sh) > 0 &&
                (best < 0 || kFfaShape[sh].stage > kFfaShape[best].stage))
                best = sh;
        if (best >= 0) {
            *shape = best;
            *cap = 0;
            most = ffa_cap_rows((int)bins, best);
        }
    }
    *cap = (*cap > 0 && *cap < most) ? *cap : most;
}

// Workgroups of a shape resident at once: as many per CU as its LDS allows.
int shape_slots(int cus, int shape)
{
    return std::max(1, cus) * std::max(1, (160 * 1024) / (kFfaShape[shape].lds_bytes + 1024));
}

struct Group {
    int s0 = 0, s1 = 0;            // series [s0, s1)
    int tr0 = 0, ntr = 0;          // transforms in the device table
    std::vector<Run> runs;         // [tr0, tr0 + ntr) split by kernel variant
    int ds0 = 0;                   // DsSeriesDev entries [ds0, ds0 + s1 - s0)
    int total_chunks = 0;
    int halo = 2;
    // series with more packed steps than one launch holds (kDsMaxSteps):
    // their steps [k * kDsMaxSteps, (k + 1) * kDsMaxSteps) in launch k >= 1,
    // DsSeriesDev entries [ds0, ds0 + count) of those series; every launch
    // of the group stages its chunks with the same halo
    struct DsLaunch {
        int ds0, count, total_chunks;
    };
    std::vector<DsLaunch> more_ds;
    int nds_series = 0;            // series with at least one packed step
    // steps whose factor exceeds the packed kernel's limit (the longest
    // periods at few bins): one direct downsampling launch each
    struct DirectStep {
        int series;
        double f;
        int64_t n, out_offset;
    };
    std::vector<DirectStep> direct;
};

}  // namespace

// The tiling table in effect, 7 values per class (what rt_batch_set_tiling
// took): batches created under equal tables tile alike.
std::vector<int32_t> batch_tiling_flat()
{
    std::vector<int32_t> out;
    for (const TileClass& c : g_tile_classes)
        out.insert(out.end(), {c.rows_lo, c.rows_hi, c.bins_lo, c.bins_hi, c.shape, c.cap, c.lmax});
    return out;
}

struct Batch {
    int device = 0;
    int64_t nseries = 0;
    double pmin = 0, pmax = 0;
    // per series
    std::vector<int64_t> N;
    std::vector<double> tsamp;
    std::vector<int64_t> bmin, bmax;
    std::vector<int> plan_of;            // index into plans
    std::vector<int> wgroup_of;          // index into width groups
    std::vector<int64_t> ds_off;         // offset in the group ds buffer
    std::vector<int64_t> prep_off;       // offset in the group prepared buffer
    std::vector<SeriesPlan> plans;
    // width groups: (bmin, bmax, widths) -> tables
    std::vector<std::vector<uint64_t>> wlists;
    std::vector<int64_t> wg_bmin, wg_bmax;
    std::vector<int64_t> wg_hb_off, wg_w_off;
    // options
    int deredden = 0, normalise = 0;
    double rmed_width = 4.0;
    int64_t rmed_minpts = 101;
    // groups
    std::vector<Group> groups;
    std::vector<FfaTransform> host_trs;  // all groups, src/snr filled at bind
    std::vector<int> tr_series;
    std::vector<int> tr_plan_index;      // transform index within its plan
    std::vector<int64_t> tr_rec_off;     // per transform: word offset of its records, or -1
    std::vector<int64_t> tr_task_off;    // per transform: first entry of its task list, or -1
    std::vector<FfaTaskDesc> task_descs; // task lists of the transforms several series share
    std::vector<FfaRecordSpec> rec_specs;  // tile records shared by equal transforms
    int64_t rec_words_total = 0;
    std::vector<DsSeriesDev> host_ds;
    std::vector<int> ds_series;            // per host_ds entry: its series
    std::vector<DsStepDev> host_steps;
    std::vector<PrepSeriesDev> host_prep;  // per series, pointers filled at bind
    std::vector<int64_t> work_off;         // per series, offset in the group work buffer
    // device memory
    float* d_tables = nullptr;
    FfaTransform* d_trs = nullptr;
    DsSeriesDev* d_ds_series = nullptr;
    DsStepDev* d_ds_steps = nullptr;
    int32_t* d_rec = nullptr;
    FfaRecordSpec* d_rec_specs = nullptr;
    FfaTaskDesc* d_task_descs = nullptr;
    PrepSeriesDev* d_prep_series = nullptr;
    float* d_ds = nullptr;
    float* d_prep = nullptr;
    float* d_scratch = nullptr;
    float* d_work = nullptr;             // dereddening + normalisation scratch
    int* d_queue = nullptr;        // one counter per run
    int nqueues = 0;
    int* d_err = nullptr;
    int64_t ds_floats = 0, prep_floats = 0, slot_floats = 0, work_floats = 0;
    int slots = 0;
    int cus = 1;
    int shape = 0;
    // bound buffers
    std::vector<const float*> inputs;
    std::vector<float*> outputs;
    bool bound = false;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

static void free_batch(Batch* b)
{
    if (b->graph_stream) hipStreamDestroy(b->graph_stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (void* p : {(void*)b->d_rec, (void*)b->d_rec_specs, (void*)b->d_task_descs, (void*)b->d_tables, (void*)b->d_trs, (void*)b->d_ds_series,
                    (void*)b->d_ds_steps, (void*)b->d_prep_series, (void*)b->d_ds,
                    (void*)b->d_prep,
                    (void*)b->d_scratch, (void*)b->d_work, (void*)b->d_queue, (void*)b->d_err})
        if (p) hipFree(p);
    delete b;
}

static int64_t transform_cost(const Transform& t)
{
    return (int64_t)t.rows * t.bins * (ffa_height(t.rows) + 4);
}

// Tile records of one transform shape (FfaRecord layout in tasks.h): one per
// (pass, node size class, tile) in the order next_ffa_task indexes them;
// appends their specs to b->rec_specs and sets f->rec_first.  Returns the
// word offset of the first record.
static int64_t plan_records(Batch* b, FfaTransform* f)
{
    const int64_t stride = rec_words(f->cap);
    const int64_t off0 = b->rec_words_total;
    int32_t count = 0;
    auto add = [&](int32_t m, int32_t lo, int32_t hi, int32_t levels, int32_t cone) {
        FfaRecordSpec r = {};
        r.off = b->rec_words_total;
        r.bins = f->bins;
        r.cap = f->cap;
        r.node_m = m;
        r.lo = lo;
        r.hi = hi;
        r.levels = levels;
        r.cone = cone;
        b->rec_specs.push_back(r);
        b->rec_words_total += stride;
        ++count;
    };
    std::memset(f->rec_first, 0, sizeof(f->rec_first));
    if (f->npass == 1) {
        add(f->rows, 0, f->rows, f->levels0, 0);
        return off0;
    }
    for (int p = 0; p < f->npass; ++p) {
        const int d = f->depth[p];
        const int L = p > 0 ? f->depth[p - 1] - d : 0;
        const bool last = p == f->npass - 1;
        const int K = p == 0 ? 0 : last ? f->k_last : f->cap - 2 * ((1 << L) - 1);
        for (int c = 0; c < 2; ++c) {
            const int m = (f->rows >> d) + c;
            f->rec_first[p][c] = count;
            if (m < 1) continue;
            if (p == 0) {
                add(m, 0, m, ffa_height(m), 0);
            } else {
                for (int lo = 0; lo < m; lo += K) add(m, lo, std::min(lo + K, m), L, 1);
            }
        }
    }
    return off0;
}

// The task list of one transform (its records, if any, planned): the tasks
// next_ffa_task gives, as descriptors.  Returns the first entry, or -1 if a
// task does not fit the descriptor's fields.
static int64_t plan_task_list(Batch* b, const FfaTransform& f, int32_t* ntasks)
{
    // stand-in addresses: the enumeration only compares and offsets them
    static const float kInput[1] = {0}, kBufA[1] = {0}, kBufB[1] = {0};
    static const int32_t kRec[1] = {0};
    FfaTransform tr = f;
    tr.src = kInput;
    tr.rec = kRec;
    float* buf_a = const_cast<float*>(kBufA);
    float* buf_b = const_cast<float*>(kBufB);
    const int64_t first = (int64_t)b->task_descs.size();
    FfaTask tk = {};
    int p = 0, i = 0, lo = 0, j = 0, node = -1, r0 = 0, nm = 0;
    while (next_ffa_task(tr, tr.depth, &tr.rec_first[0][0], p, i, lo, j, node, r0, nm, buf_a,
                         buf_b, tk)) {
        FfaTaskDesc d = {};
        d.node_r0 = tk.node_r0;
        d.node_m = tk.node_m;
        d.lo_rows = tk.lo | ((tk.hi - tk.lo) << 16);
        const int src = tk.src == buf_a ? 1 : tk.src == buf_b ? 2 : 0;
        const int dst = tk.dst == buf_b ? 1 : 0;
        const int64_t rec = (tk.rec - kRec) / rec_words(f.cap);
        if (rec >= (1 << 19) || tk.levels > 15 || tk.lo >= (1 << 16) || tk.hi - tk.lo >= (1 << 15)) {
            b->task_descs.resize((size_t)first);
            *ntasks = 0;
            return -1;
        }
        d.flags = tk.levels | (tk.cone << 4) | (tk.kind << 5) | (src << 6) | (dst << 8) |
                  (int32_t)(rec << 12);
        b->task_descs.push_back(d);
    }
    *ntasks = (int32_t)((int64_t)b->task_descs.size() - first);
    return first;
}

static int scrunch_factor(int64_t width_samples, int64_t minpts)
{
    return (int)std::max(1.0, width_samples / (double)minpts);
}

}  // namespace rt

using namespace rt;

extern "C" {

int rt_batch_create(int64_t nseries, const int64_t* sizes, const double* tsamps,
                    const uint64_t* widths_flat, const int64_t* widths_count, double period_min,
                    double period_max, const int64_t* bins_min, const int64_t* bins_max,
                    int deredden, double rmed_width, int64_t rmed_minpts, int normalise,
                    int64_t group_size, void** out)
{
    if (nseries < 1) return bfail(RT_EINVAL, "batch needs at least one series");
    if (deredden && !(rmed_minpts % 2)) return bfail(RT_EINVAL, "min_points must be an odd number");
    Batch* b = new Batch();
    BT_HIP(hipGetDevice(&b->device));
    b->nseries = nseries;
    b->pmin = period_min;
    b->pmax = period_max;
    b->deredden = deredden;
    b->normalise = normalise;
    b->rmed_width = rmed_width;
    b->rmed_minpts = rmed_minpts;
    b->shape = current_ffa_shape();

    // plans (shared by identical search parameters) and width groups
    std::map<std::tuple<int64_t, double, int64_t, int64_t>, int> plan_ids;
    std::map<std::tuple<int64_t, int64_t, std::vector<uint64_t>>, int> wg_ids;
    const uint64_t* wp = widths_flat;
    for (int64_t i = 0; i < nseries; ++i) {
        b->N.push_back(sizes[i]);
        b->tsamp.push_back(tsamps[i]);
        b->bmin.push_back(bins_min[i]);
        b->bmax.push_back(bins_max[i]);
        const auto key = std::make_tuple(sizes[i], tsamps[i], bins_min[i], bins_max[i]);
        auto it = plan_ids.find(key);
        if (it == plan_ids.end()) {
            SeriesPlan p;
            const std::string e = plan_series(sizes[i], tsamps[i], period_min, period_max,
                                              bins_min[i], bins_max[i], &p);
            if (!e.empty()) {
                free_batch(b);
                return bfail(RT_EINVAL, e);
            }
            it = plan_ids.emplace(key, (int)b->plans.size()).first;
            b->plans.push_back(std::move(p));
        }
        b->plan_of.push_back(it->second);
        std::vector<uint64_t> w(wp, wp + widths_count[i]);
        wp += widths_count[i];
        for (uint64_t x : w)
            if (!(x > 0 && x < (uint64_t)bins_min[i])) {
                free_batch(b);
                return bfail(RT_EINVAL, "trial widths must be all > 0 and < columns");
            }
        const auto wkey = std::make_tuple(bins_min[i], bins_max[i], w);
        auto wt = wg_ids.find(wkey);
        if (wt == wg_ids.end()) {
            wt = wg_ids.emplace(wkey, (int)b->wlists.size()).first;
            b->wlists.push_back(w);
            b->wg_bmin.push_back(bins_min[i]);
            b->wg_bmax.push_back(bins_max[i]);
        }
        b->wgroup_of.push_back(wt->second);
        if (deredden) {
            const int64_t ws = (int64_t)std::llround(rmed_width / tsamps[i]);
            const int sf = scrunch_factor(ws, rmed_minpts);
            const int64_t nl = sizes[i] / sf;
            if (sf == 1 ? !(ws % 2 && ws < sizes[i]) : !(rmed_minpts < nl)) {
                free_batch(b);
                return bfail(RT_EINVAL, "running median width must be odd and < size");
            }
        }
    }

    // S/N tables of every width group: (h + b, b) per bins, then widths
    std::vector<float> tables;
    for (size_t g = 0; g < b->wlists.size(); ++g) {
        const auto& w = b->wlists[g];
        b->wg_hb_off.push_back((int64_t)tables.size());
        for (int64_t bins = b->wg_bmin[g]; bins <= b->wg_bmax[g]; ++bins)
            for (uint64_t x : w) {
                float hpb, bb;
                boxcar_hb(bins, (int64_t)x, &hpb, &bb);
                tables.push_back(hpb);
                tables.push_back(bb);
            }
        b->wg_w_off.push_back((int64_t)tables.size());
        for (uint64_t x : w) {
            const int32_t v = (int32_t)x;
            float f;
            std::memcpy(&f, &v, sizeof(f));
            tables.push_back(f);
        }
    }

    // series per plan: the transforms of a plan that several series share
    // get tile records, built once per run and read by every one of them
    std::vector<int> plan_series_count(b->plans.size(), 0);
    for (int64_t i = 0; i < nseries; ++i) ++plan_series_count[b->plan_of[i]];
    // (plan, transform) -> (record offset, task list offset, transform with its record offsets)
    std::map<std::pair<int, int>, std::tuple<int64_t, int64_t, FfaTransform>> rec_of;
    // task lists of transforms without records, shared by every transform of
    // the same tiling (rows, rows evaluated, tile rows, pass depths), whatever
    // its series or bins: (key) -> (task list offset, tasks)
    std::map<std::vector<int32_t>, std::pair<int64_t, int32_t>> list_of;

    // groups of series; per group: packed downsampling descriptors and the
    // transform queue sorted by decreasing cost
    if (group_size < 1) group_size = 64;
    int64_t max_ds = 1, max_prep = 1, max_fold = 1, max_work = 1;
    b->ds_off.assign(nseries, 0);
    b->prep_off.assign(nseries, 0);
    for (int64_t s0 = 0; s0 < nseries; s0 += group_size) {
        Group g;
        g.s0 = (int)s0;
        g.s1 = (int)std::min<int64_t>(nseries, s0 + group_size);
        g.ds0 = (int)b->host_ds.size();
        int64_t ds_run = 0, prep_run = 0;
        double fmax = 1.0;
        std::vector<std::pair<int64_t, int>> order;  // (cost, transform index in host_trs)
        const int tr_base = (int)b->host_trs.size();
        std::vector<FfaTransform> local;
        std::vector<int> local_series, local_pidx;
        std::vector<int64_t> local_rec, local_task;
        std::vector<int> packed_steps;  // per series of the group
        for (int s = g.s0; s < g.s1; ++s) {
            const SeriesPlan& p = b->plans[b->plan_of[s]];
            b->ds_off[s] = ds_run;
            b->prep_off[s] = prep_run;
            prep_run += p.size;
            DsSeriesDev d = {};
            d.N = p.size;
            d.step0 = (int32_t)b->host_steps.size();
            d.chunk_offset = g.total_chunks;
            int nst = 0;
            for (const Step& st : p.steps) {
                if (st.ds_offset < 0 || st.n == 0) continue;
                if (st.f > kDsPackedMaxF) {
                    g.direct.push_back({s, st.f, st.n, st.ds_offset});
                    continue;
                }
                DsStepDev ds;
                split_double(st.f, &ds.mf, &ds.ef);
                ds.inv_f = (float)(1.0 / st.f);
                ds.n = st.n;
                ds.out_offset = st.ds_offset;  // relative to the series' region
                b->host_steps.push_back(ds);
                fmax = std::max(fmax, st.f);
                ++nst;
            }
            d.nsteps = std::min(nst, kDsMaxSteps);
            packed_steps.push_back(nst);
            g.total_chunks += (int)((p.size + kDsChunk - 1) / kDsChunk);
            b->host_ds.push_back(d);
            b->ds_series.push_back(s);
            ds_run += p.ds_total;
            const int wg = b->wgroup_of[s];
            for (size_t it = 0; it < p.tr.size(); ++it) {
                const Transform& t = p.tr[it];
                if (t.rows_eval <= 0) continue;
                FfaTransform f = {};
                f.bins = t.bins;
                f.rows = t.rows;
                f.rows_eval = t.rows_eval;
                f.nw = (int32_t)b->wlists[wg].size();
                f.stdnoise = t.stdnoise;
                int tshape = 0, tcap = 0, tlmax = 0;
                tile_for(t.rows, t.bins, b->shape, &tshape, &tcap, &tlmax);
                f.cap = tcap;
                f.shape = tshape;
                f.series = s;
                f.wmax = 0;
                for (uint64_t x : b->wlists[wg]) f.wmax = std::max<int32_t>(f.wmax, (int32_t)x);
                f.ext_rows = ffa_snr_ext_rows(t.bins, f.wmax, f.nw, tshape);
                f.k_last = 0;
                if (t.rows <= f.cap) {
                    f.npass = 1;
                    f.levels0 = ffa_height(t.rows);
                } else {
                    const int np = count_transform_passes(t.rows, t.bins, f.cap, tlmax, f.ext_rows);
                    if (np < 0 || np > kMaxPasses) {
                        free_batch(b);
                        return bfail(RT_EINVAL, "bins too large for the FFA tiling");
                    }
                    std::vector<int> depths = transform_pass_depths(t.rows, f.cap, tlmax, f.ext_rows);
                    f.npass = (int32_t)depths.size();
                    for (size_t q = 0; q < depths.size(); ++q) f.depth[q] = depths[q];
                    f.k_last = (int32_t)last_tile_rows(
                        f.cap, depths[depths.size() - 2] - depths[depths.size() - 1], f.ext_rows);
                    max_fold = std::max<int64_t>(max_fold, (int64_t)t.rows * t.bins);
                }
                int64_t rec_off = -1, task_off = -1;
                if (plan_series_count[b->plan_of[s]] < kRecMinSeries) {
                    std::vector<int32_t> key = {f.rows, f.rows_eval, f.cap, f.npass, f.levels0,
                                                f.k_last};
                    for (int q = 0; q < f.npass && q < kMaxPasses; ++q) key.push_back(f.depth[q]);
                    auto lt = list_of.find(key);
                    if (lt == list_of.end()) {
                        int32_t n = 0;
                        const int64_t off = plan_task_list(b, f, &n);
                        lt = list_of.emplace(std::move(key), std::make_pair(off, n)).first;
                    }
                    task_off = lt->second.first;
                    f.ntasks = lt->second.second;
                } else {
                    const auto key = std::make_pair(b->plan_of[s], (int)it);
                    auto rt = rec_of.find(key);
                    if (rt == rec_of.end()) {
                        FfaTransform g = f;
                        // records up to a memory budget; past it the tasks
                        // of further shapes build their tables themselves
                        const int64_t off = b->rec_words_total * 4 < kRecMaxBytes
                                                ? plan_records(b, &g)
                                                : -1;
                        const int64_t toff = plan_task_list(b, g, &g.ntasks);
                        rt = rec_of.emplace(key, std::make_tuple(off, toff, g)).first;
                    }
                    rec_off = std::get<0>(rt->second);
                    task_off = std::get<1>(rt->second);
                    const FfaTransform& g = std::get<2>(rt->second);
                    std::memcpy(f.rec_first, g.rec_first, sizeof(f.rec_first));
                    f.ntasks = g.ntasks;
                }
                if (task_off < 0) {
                    free_batch(b);
                    return bfail(RT_EINVAL, "transform too tall for the batch task list");
                }
                order.emplace_back(transform_cost(t), (int)local.size());
                local.push_back(f);
                local_series.push_back(s);
                local_pidx.push_back((int)it);
                local_rec.push_back(rec_off);
                local_task.push_back(task_off);
            }
        }
        // by kernel variant and workgroup shape, then largest first
        std::stable_sort(order.begin(), order.end(),
                         [&local](const std::pair<int64_t, int>& a, const std::pair<int64_t, int>& c) {
                             const int va = ffa_variant(local[a.second].bins);
                             const int vc = ffa_variant(local[c.second].bins);
                             if (va != vc) return va < vc;
                             if (local[a.second].shape != local[c.second].shape)
                                 return local[a.second].shape < local[c.second].shape;
                             return a.first > c.first;
                         });
        for (const auto& o : order) {
            const int v = ffa_variant(local[o.second].bins);
            const int sh = local[o.second].shape;
            if (g.runs.empty() || g.runs.back().variant != v || g.runs.back().shape != sh) {
                Run r;
                r.tr0 = (int)b->host_trs.size();
                r.variant = v;
                r.shape = sh;
                r.queue = b->nqueues++;
                g.runs.push_back(r);
            }
            ++g.runs.back().ntr;
            b->host_trs.push_back(local[o.second]);
            b->tr_series.push_back(local_series[o.second]);
            b->tr_plan_index.push_back(local_pidx[o.second]);
            b->tr_rec_off.push_back(local_rec[o.second]);
            b->tr_task_off.push_back(local_task[o.second]);
        }
        g.tr0 = tr_base;
        g.ntr = (int)order.size();
        g.halo = (int)std::ceil(fmax) + 2;
        for (int k = 1;; ++k) {
            Group::DsLaunch l = {(int)b->host_ds.size(), 0, 0};
            for (int i = 0; i < g.s1 - g.s0; ++i) {
                DsSeriesDev d = b->host_ds[(size_t)g.ds0 + i];
                const int total = packed_steps[(size_t)i];
                if (total <= k * kDsMaxSteps) continue;
                d.step0 += k * kDsMaxSteps;
                d.nsteps = std::min(kDsMaxSteps, total - k * kDsMaxSteps);
                d.chunk_offset = l.total_chunks;
                l.total_chunks += (int)((d.N + kDsChunk - 1) / kDsChunk);
                b->host_ds.push_back(d);
                b->ds_series.push_back(g.s0 + i);
                ++l.count;
            }
            if (!l.count) break;
            g.more_ds.push_back(l);
        }
        max_ds = std::max(max_ds, ds_run);
        max_prep = std::max(max_prep, prep_run);
        b->groups.push_back(g);
    }
    // preparation work areas of every series of a group, side by side:
    // block means, running median, normalisation partials and moments
    b->host_prep.assign(nseries, PrepSeriesDev{});
    b->work_off.assign(nseries, 0);
    for (const Group& g : b->groups) {
        int64_t run = 0;
        for (int s = g.s0; s < g.s1; ++s) {
            PrepSeriesDev& d = b->host_prep[s];
            d.N = sizes[s];
            d.sf = 1;
            d.nl = sizes[s];
            d.width = 1;
            d.dereddened = deredden;
            b->work_off[s] = run;
            if (deredden) {
                const int64_t ws = (int64_t)std::llround(rmed_width / tsamps[s]);
                d.sf = scrunch_factor(ws, rmed_minpts);
                d.nl = sizes[s] / d.sf;
                d.width = (int32_t)(d.sf == 1 ? ws : rmed_minpts);
                run += (d.sf == 1 ? 0 : d.nl) + d.nl;
            }
            run += 2 * (kNormaliseParts + 2);
        }
        max_work = std::max(max_work, run);
    }

    int cus = 0;
    BT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    // scratch slots: the most workgroups any run's shape keeps resident
    b->cus = std::max(1, cus);
    b->slots = 1;
    for (const Group& g : b->groups)
        for (const Run& r : g.runs) b->slots = std::max(b->slots, shape_slots(cus, r.shape));
    b->ds_floats = max_ds;
    b->prep_floats = (deredden || normalise) ? max_prep : 1;
    b->slot_floats = 2 * max_fold;
    b->work_floats = max_work;

    BT_HIP(hipMalloc(&b->d_tables, std::max<size_t>(tables.size(), 1) * sizeof(float)));
    if (!tables.empty())
        BT_HIP(hipMemcpy(b->d_tables, tables.data(), tables.size() * sizeof(float),
                         hipMemcpyHostToDevice));
    BT_HIP(hipMalloc(&b->d_trs, std::max<size_t>(b->host_trs.size(), 1) * sizeof(FfaTransform)));
    BT_HIP(hipMalloc(&b->d_ds_series, std::max<size_t>(b->host_ds.size(), 1) * sizeof(DsSeriesDev)));
    BT_HIP(hipMalloc(&b->d_ds_steps, std::max<size_t>(b->host_steps.size(), 1) * sizeof(DsStepDev)));
    BT_HIP(hipMalloc(&b->d_prep_series, std::max<size_t>(b->host_prep.size(), 1) * sizeof(PrepSeriesDev)));
    BT_HIP(hipMalloc(&b->d_ds, (size_t)b->ds_floats * sizeof(float)));
    BT_HIP(hipMalloc(&b->d_prep, (size_t)b->prep_floats * sizeof(float)));
    BT_HIP(hipMalloc(&b->d_scratch, (size_t)b->slots * b->slot_floats * sizeof(float)));
    BT_HIP(hipMalloc(&b->d_work, (size_t)b->work_floats * sizeof(float)));
    BT_HIP(hipMalloc(&b->d_queue, std::max<size_t>(b->nqueues, 1) * sizeof(int)));
    BT_HIP(hipMalloc(&b->d_err, sizeof(int)));
    BT_HIP(hipMemset(b->d_err, 0, sizeof(int)));
    if (!b->rec_specs.empty()) {
        BT_HIP(hipMalloc(&b->d_rec, (size_t)b->rec_words_total * sizeof(int32_t)));
        BT_HIP(hipMemset(b->d_rec, 0, (size_t)b->rec_words_total * sizeof(int32_t)));
        BT_HIP(hipMalloc(&b->d_rec_specs, b->rec_specs.size() * sizeof(FfaRecordSpec)));
        BT_HIP(hipMemcpy(b->d_rec_specs, b->rec_specs.data(),
                         b->rec_specs.size() * sizeof(FfaRecordSpec), hipMemcpyHostToDevice));
    }
    if (!b->task_descs.empty()) {
        BT_HIP(hipMalloc(&b->d_task_descs, b->task_descs.size() * sizeof(FfaTaskDesc)));
        BT_HIP(hipMemcpy(b->d_task_descs, b->task_descs.data(),
                         b->task_descs.size() * sizeof(FfaTaskDesc), hipMemcpyHostToDevice));
    }
    for (size_t k = 0; k < b->host_trs.size(); ++k) {
        b->host_trs[k].rec = b->tr_rec_off[k] >= 0 ? b->d_rec + b->tr_rec_off[k] : nullptr;
        b->host_trs[k].tasks =
            b->tr_task_off[k] >= 0 ? b->d_task_descs + b->tr_task_off[k] : nullptr;
    }
    *out = b;
    return RT_OK;
}

int rt_batch_length(void* handle, int64_t i, int64_t* length)
{
    Batch* b = (Batch*)handle;
    if (i < 0 || i >= b->nseries) return bfail(RT_EINVAL, "series index out of range");
    *length = b->plans[b->plan_of[i]].length;
    return RT_OK;
}

int rt_batch_grid(void* handle, int64_t i, double* h_periods, uint32_t* h_foldbins)
{
    Batch* b = (Batch*)handle;
    if (i < 0 || i >= b->nseries) return bfail(RT_EINVAL, "series index out of range");
    plan_periods(b->plans[b->plan_of[i]], h_periods, h_foldbins);
    return RT_OK;
}

int rt_batch_bind(void* handle, const float* const* h_inputs, float* const* h_snrs)
{
    Batch* b = (Batch*)handle;
    b->inputs.assign(h_inputs, h_inputs + b->nseries);
    b->outputs.assign(h_snrs, h_snrs + b->nseries);
    const bool prep = b->deredden || b->normalise;
    for (size_t k = 0; k < b->host_ds.size(); ++k) {
        const int s = b->ds_series[k];
        b->host_ds[k].x = prep ? b->d_prep + b->prep_off[s] : b->inputs[s];
        b->host_ds[k].out = b->d_ds + b->ds_off[s];
    }
    for (int64_t i = 0; i < b->nseries; ++i) {
        PrepSeriesDev& d = b->host_prep[i];
        float* w = b->d_work + b->work_off[i];
        d.x = b->inputs[i];
        d.out = b->d_prep + b->prep_off[i];
        d.lores = w;
        d.rmed = (b->deredden && d.sf > 1) ? w + d.nl : w;
        d.stats = w + (b->deredden ? (d.sf == 1 ? 0 : d.nl) + d.nl : 0);
    }
    for (size_t k = 0; k < b->host_trs.size(); ++k) {
        FfaTransform& f = b->host_trs[k];
        const int s = b->tr_series[k];
        const SeriesPlan& p = b->plans[b->plan_of[s]];
        const Transform& t = p.tr[b->tr_plan_index[k]];
        const Step& st = p.steps[t.step];
        const int wg = b->wgroup_of[s];
        const float* base = prep ? b->d_prep + b->prep_off[s] : b->inputs[s];
        f.src = (st.ds_offset < 0) ? base : b->d_ds + b->ds_off[s] + st.ds_offset;
        f.snr = b->outputs[s] + t.out_row * f.nw;
        f.widths = (const int32_t*)(b->d_tables + b->wg_w_off[wg]);
        f.hb = b->d_tables + b->wg_hb_off[wg] + (t.bins - b->wg_bmin[wg]) * 2 * f.nw;
    }
    BT_HIP(hipMemcpy(b->d_trs, b->host_trs.data(), b->host_trs.size() * sizeof(FfaTransform),
                     hipMemcpyHostToDevice));
    BT_HIP(hipMemcpy(b->d_ds_series, b->host_ds.data(), b->host_ds.size() * sizeof(DsSeriesDev),
                     hipMemcpyHostToDevice));
    BT_HIP(hipMemcpy(b->d_ds_steps, b->host_steps.data(),
                     b->host_steps.size() * sizeof(DsStepDev), hipMemcpyHostToDevice));
    BT_HIP(hipMemcpy(b->d_prep_series, b->host_prep.data(),
                     b->host_prep.size() * sizeof(PrepSeriesDev), hipMemcpyHostToDevice));
    b->bound = true;
    if (b->exec) {
        hipGraphExecDestroy(b->exec);
        hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    return RT_OK;
}

// Optional per-phase event timing of an enqueue (events on the same stream).
struct PhaseTimer {
    std::vector<hipEvent_t> ev;  // 4 per group: start, after prep, after ds, after ffa
};

static int enqueue_batch(Batch* b, hipStream_t s, PhaseTimer* pt = nullptr)
{
    // every run starts with a clear error flag (a fill kernel, not a memset
    // node, so that the run can be captured into a graph)
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, b->d_err, 1, 0);
    BT_HIP(hipGetLastError());
    // the tile records of the transforms that several series share
    if (!b->rec_specs.empty())
        BT_HIP(launch_ffa_records(b->d_rec_specs, (int)b->rec_specs.size(), b->d_rec, s));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const Group& g = b->groups[gi];
        if (pt) BT_HIP(hipEventRecord(pt->ev[4 * gi + 0], s));
        // prepare: dereddening then normalisation of every series of the
        // group in one set of launches
        if (b->deredden || b->normalise) {
            int64_t max_n = 1, max_lores = 0;
            int max_width = 1;
            for (int i = g.s0; i < g.s1; ++i) {
                const PrepSeriesDev& d = b->host_prep[i];
                max_n = std::max(max_n, d.N);
                max_lores = std::max(max_lores, d.sf > 1 ? d.nl : d.N);
                max_width = std::max(max_width, (int)d.width);
            }
            BT_HIP(launch_prep_batch(b->d_prep_series + g.s0, g.s1 - g.s0, max_n, max_lores,
                                     max_width, b->deredden, b->normalise, s));
        }
        if (pt) BT_HIP(hipEventRecord(pt->ev[4 * gi + 1], s));
        if (g.total_chunks > 0)
            BT_HIP(launch_downsample_packed(b->d_ds_series + g.ds0, g.s1 - g.s0, b->d_ds_steps,
                                            g.total_chunks, kDsChunk, g.halo, s));
        for (const Group::DsLaunch& l : g.more_ds)
            BT_HIP(launch_downsample_packed(b->d_ds_series + l.ds0, l.count, b->d_ds_steps,
                                            l.total_chunks, kDsChunk, g.halo, s));
        for (const Group::DirectStep& d : g.direct) {
            const float* x = (b->deredden || b->normalise) ? b->d_prep + b->prep_off[d.series]
                                                            : b->inputs[d.series];
            BT_HIP(launch_downsample_direct(x, b->N[d.series], d.f, d.n,
                                            b->d_ds + b->ds_off[d.series] + d.out_offset, s));
        }
        if (!g.runs.empty()) {
            // queue counters of the group's runs are consecutive
            hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s,
                               b->d_queue + g.runs[0].queue, (int)g.runs.size(), 0);
            BT_HIP(hipGetLastError());
        }
        if (pt) BT_HIP(hipEventRecord(pt->ev[4 * gi + 2], s));
        for (const Run& r : g.runs)
            BT_HIP(launch_ffa_batch(b->d_trs + r.tr0, r.ntr, b->d_queue + r.queue, b->d_scratch,
                                    b->slot_floats, shape_slots(b->cus, r.shape), b->d_err, s,
                                    r.shape, r.variant));
        if (pt) BT_HIP(hipEventRecord(pt->ev[4 * gi + 3], s));
    }
    return RT_OK;
}

int rt_batch_run(void* handle, int use_graph, void* stream)
{
    Batch* b = (Batch*)handle;
    if (!b->bound) return bfail(RT_EINVAL, "rt_batch_bind must be called before rt_batch_run");
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) return enqueue_batch(b, s);
    if (!b->exec) {
        // capture on a private stream (the legacy default stream cannot be
        // captured); the graph is then launched on the caller's stream
        if (!b->graph_stream) BT_HIP(hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking));
        hipStream_t cs = b->graph_stream;
        BT_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int st = enqueue_batch(b, cs);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &graph);
        if (st) {
            if (graph) hipGraphDestroy(graph);
            return st;
        }
        BT_HIP(e);
        b->graph = graph;
        BT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    BT_HIP(hipGraphLaunch(b->exec, s));
    return RT_OK;
}

int rt_batch_run_timed(void* handle, void* stream, double* ms)
{
    // ms[0] preparation, ms[1] downsampling, ms[2] FFA + S/N, ms[3] FFA launches (one per
    // kernel variant per group)
    Batch* b = (Batch*)handle;
    if (!b->bound) return bfail(RT_EINVAL, "rt_batch_bind must be called before rt_batch_run");
    hipStream_t s = (hipStream_t)stream;
    PhaseTimer pt;
    pt.ev.resize(4 * b->groups.size());
    for (auto& e : pt.ev) BT_HIP(hipEventCreate(&e));
    int st = enqueue_batch(b, s, &pt);
    if (st == RT_OK) {
        BT_HIP(hipStreamSynchronize(s));
        double acc[3] = {0, 0, 0};
        for (size_t gi = 0; gi < b->groups.size(); ++gi)
            for (int k = 0; k < 3; ++k) {
                float t = 0;
                BT_HIP(hipEventElapsedTime(&t, pt.ev[4 * gi + k], pt.ev[4 * gi + k + 1]));
                acc[k] += t;
            }
        ms[0] = acc[0];
        ms[1] = acc[1];
        ms[2] = acc[2];
        size_t runs = 0;
        for (const Group& g : b->groups) runs += g.runs.size();
        ms[3] = (double)runs;
    }
    for (auto& e : pt.ev) hipEventDestroy(e);
    return st;
}

int rt_batch_check(void* handle, void* stream)
{
    Batch* b = (Batch*)handle;
    int herr = 0;
    BT_HIP(hipStreamSynchronize((hipStream_t)stream));
    BT_HIP(hipMemcpy(&herr, b->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & 1) return bfail(RT_EDEVICE, "FFA tile exceeded its LDS budget");
    if (herr) return bfail(RT_EDEVICE, "FFA transform launched with the wrong kernel variant");
    return RT_OK;
}

int rt_batch_set_tiling(const int32_t* entries, int n)
{
    if (n < 0 || (n > 0 && entries == nullptr)) return bfail(RT_EINVAL, "bad tiling table");
    std::vector<TileClass> t;
    for (int i = 0; i < n; ++i) {
        const int32_t* e = entries + 7 * i;
        TileClass c = {e[0], e[1], e[2], e[3], e[4], e[5], e[6]};
        if (c.rows_lo < 0 || c.rows_hi <= c.rows_lo || c.bins_lo < 1 || c.bins_hi <= c.bins_lo ||
            c.shape < 0 || c.shape >= kFfaShapes || c.cap < 0 || c.lmax < 0 ||
            c.lmax > kMaxConeLevels)
            return bfail(RT_EINVAL, "bad tiling table entry");
        t.push_back(c);
    }
    g_tile_classes = t;
    return RT_OK;
}

int rt_batch_stats(void* handle, int64_t* stats)
{
    // [0] transforms, [1] groups, [2] slots, [3] scratch bytes, [4] ds bytes,
    // [5] fold input floats (sum rows*bins), [6] total periods, [7] FFA adds,
    // [8] tile records, [9] tile record bytes
    Batch* b = (Batch*)handle;
    int64_t fold = 0, periods = 0, adds = 0;
    for (int64_t i = 0; i < b->nseries; ++i) {
        const SeriesPlan& p = b->plans[b->plan_of[i]];
        periods += p.length;
        for (const Transform& t : p.tr) {
            fold += (int64_t)t.rows * t.bins;
            adds += ffa_adds(t.rows) * t.bins;
        }
    }
    stats[0] = (int64_t)b->host_trs.size();
    stats[1] = (int64_t)b->groups.size();
    stats[2] = b->slots;
    stats[3] = (int64_t)b->slots * b->slot_floats * 4;
    stats[4] = b->ds_floats * 4;
    stats[5] = fold;
    stats[6] = periods;
    stats[7] = adds;
    stats[8] = (int64_t)b->rec_specs.size();
    stats[9] = b->rec_words_total * 4;
    return RT_OK;
}

int rt_batch_destroy(void* handle)
{
    if (handle) free_batch((Batch*)handle);
    return RT_OK;
}

}  // extern "C"
