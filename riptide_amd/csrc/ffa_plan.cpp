#include "ffa_plan.h"

#include <algorithm>
#include <cmath>

namespace rt {

static constexpr int kSegBytes = (2 << kMaxConeLevels) * 5 * 4;

int ffa_lds_bytes(int shape) { return kFfaShape[shape].lds_bytes; }

int ffa_default_shape() { return 0; }

int ffa_height(int64_t m)
{
    int h = 0;
    while (((int64_t)1 << h) < m) ++h;
    return h;
}

int ffa_cap_rows(int bins, int shape)
{
    // LDS: rows of `bins` floats plus one packed table entry per row and level
    const int avail = kFfaShape[shape].lds_bytes - kSegBytes;
    // rows are ffa_pitch(bins) floats apart; one more row holds the -0.0f
    // row the copy rows add
    const int pitch = ffa_pitch(bins);
    const int cap = (avail - 4 * pitch) / (4 * (pitch + kFfaMaxLevels));
    const int k = (bins + kWave - 1) / kWave;
    const int stage = kFfaShape[shape].stage;
    if (k > stage || bins >= 4096) return 0;
    const int by_regs = (kFfaShape[shape].threads / kWave) * ffa_rows_per_wave(k, stage);
    return std::min(std::min(cap, by_regs), kFfaMaxCap);
}

// Nodes of the merge tree at depth d, in row order (block.hpp:30-43 split).
static void nodes_at_depth(int64_t r0, int64_t m, int d, std::vector<std::pair<int64_t, int64_t>>& out)
{
    if (d == 0 || m <= 1) {
        out.emplace_back(r0, m);
        return;
    }
    const int64_t mh = m >> 1;
    nodes_at_depth(r0, mh, d - 1, out);
    nodes_at_depth(r0 + mh, m - mh, d - 1, out);
}

// Smallest depth whose nodes all fit in `cap` rows.
static int leaf_depth(int64_t m, int cap)
{
    int d = 0;
    while (((m + ((int64_t)1 << d) - 1) >> d) > cap) ++d;
    return d;
}

// Rows of a cone tile with L levels whose bottom fits in cap rows: each
// internal segment grows by at most 2 rows per level.
static int64_t tile_rows(int cap, int L) { return (int64_t)cap - 2 * (((int64_t)1 << L) - 1); }

int ffa_snr_ext_rows(int bins, int wmax, int nw, int shape)
{
    if (wmax < 3 || nw < 1 || nw > kSnrMaxWidths || wmax >= bins) return 0;
    // finished rows start one row (the -0.0f row) into the dynamic LDS
    const int avail = kFfaShape[shape].lds_bytes / 4 - ffa_pitch(bins);
    return std::max(0, avail / (ffa_snr_pitch(bins, wmax) + nw));
}

int64_t last_tile_rows(int cap, int L, int ext_rows)
{
    const int64_t K = tile_rows(cap, L);
    if (ext_rows > 0 && ext_rows < K && 8 * (int64_t)ext_rows >= 7 * K) return ext_rows;
    return K;
}

// Chooses the depths d_0 = d1 > d_1 > ... > d_P = 0 of the cone passes, at
// most lmax levels per pass (the last pass's tiles as last_tile_rows).
static std::vector<int> cone_depths(int d1, int cap, int lmax, int ext_rows)
{
    std::vector<int> best;
    double best_cost = 1e300;
    for (int P = 1; P <= d1; ++P) {
        std::vector<int> depths;
        depths.push_back(d1);
        bool ok = true;
        double cost = 0;
        int rem = d1;
        for (int p = 0; p < P; ++p) {
            const int L = (rem + (P - p) - 1) / (P - p);
            if (L > lmax || L > kMaxConeLevels) { ok = false; break; }
            const int64_t K = p == P - 1 ? last_tile_rows(cap, L, ext_rows) : tile_rows(cap, L);
            if (K < 4) { ok = false; break; }
            cost += (double)(K + 2 * ((1 << L) - 1)) / (double)K;  // reads per output row
            rem -= L;
            depths.push_back(rem);
        }
        if (!ok) continue;
        cost += (P - 1);  // intermediate writes
        if (cost < best_cost) {
            best_cost = cost;
            best = depths;
        }
    }
    return best;
}

std::vector<int> transform_pass_depths(int64_t m, int cap, int lmax, int ext_rows)
{
    if (cap < 1) return {};
    if (m <= cap) return {0};
    return cone_depths(leaf_depth(m, cap), cap, lmax, ext_rows);
}

int64_t ffa_adds(int64_t m)
{
    if (m <= 1) return 0;
    const int64_t mh = m >> 1;
    return ffa_adds(mh) + ffa_adds(m - mh) + m;
}

int count_transform_passes(int64_t m, int bins, int cap, int lmax, int ext_rows)
{
    (void)bins;
    if (cap < 1) return -1;  // rows too wide for the shape's tile
    if (m <= cap) return 1;
    const int d1 = leaf_depth(m, cap);
    const std::vector<int> depths = cone_depths(d1, cap, lmax, ext_rows);
    if (depths.empty()) return -1;
    return (int)depths.size();  // leaf pass + (size-1) cone passes
}

static FfaTask base_task(const TransformSpec& s)
{
    FfaTask t = {};
    t.bins = s.bins;
    t.widths = s.widths;
    t.hb = s.hb;
    t.nw = s.nw;
    t.stdnoise = s.stdnoise;
    t.rows_eval = s.rows_eval;
    t.snr = s.snr;
    t.pad = s.cap;
    t.wmax = s.kind == 1 ? s.wmax : 0;
    t.ext_rows = s.kind == 1 ? ffa_snr_ext_rows(s.bins, s.wmax, s.nw, s.shape) : 0;
    return t;
}

int plan_transform_tasks(const TransformSpec& s, std::vector<std::vector<FfaTask>>& passes)
{
    auto ensure = [&](size_t n) {
        if (passes.size() < n) passes.resize(n);
    };
    if (s.cap < 1) return -1;  // rows too wide for the shape's tile
    if (s.m <= s.cap) {
        ensure(1);
        FfaTask t = base_task(s);
        t.src = s.src;
        t.dst = s.dst;
        t.node_r0 = 0;
        t.node_m = (int32_t)s.m;
        t.lo = 0;
        t.hi = (int32_t)s.m;
        t.levels = ffa_height(s.m);
        t.cone = 0;
        t.kind = s.kind;
        passes[0].push_back(t);
        return 1;
    }
    const int d1 = leaf_depth(s.m, s.cap);
    const int ext_rows = s.kind == 1 ? ffa_snr_ext_rows(s.bins, s.wmax, s.nw, s.shape) : 0;
    const std::vector<int> depths = cone_depths(d1, s.cap, kMaxConeLevels, ext_rows);
    if (depths.empty()) return -1;
    const int npass = (int)depths.size();
    ensure((size_t)npass);

    // pass 0: whole nodes at depth d1 -> buf_a
    std::vector<std::pair<int64_t, int64_t>> nodes;
    nodes_at_depth(0, s.m, d1, nodes);
    for (const auto& nd : nodes) {
        FfaTask t = base_task(s);
        t.src = s.src;
        t.dst = s.buf_a;
        t.node_r0 = (int32_t)nd.first;
        t.node_m = (int32_t)nd.second;
        t.lo = 0;
        t.hi = (int32_t)nd.second;
        t.levels = ffa_height(nd.second);
        t.cone = 0;
        t.kind = 0;
        passes[0].push_back(t);
    }
    // cone passes
    const float* src = s.buf_a;
    float* other = s.buf_b;
    for (int p = 1; p < npass; ++p) {
        const int d_in = depths[p - 1];
        const int d_out = depths[p];
        const int L = d_in - d_out;
        const bool last = (p == npass - 1);
        const int64_t K = last ? last_tile_rows(s.cap, L, ext_rows) : tile_rows(s.cap, L);
        nodes.clear();
        nodes_at_depth(0, s.m, d_out, nodes);
        for (const auto& nd : nodes) {
            for (int64_t lo = 0; lo < nd.second; lo += K) {
                FfaTask t = base_task(s);
                t.src = src;
                t.node_r0 = (int32_t)nd.first;
                t.node_m = (int32_t)nd.second;
                t.lo = (int32_t)lo;
                t.hi = (int32_t)std::min<int64_t>(lo + K, nd.second);
                t.levels = L;
                t.cone = 1;
                if (last) {
                    t.kind = s.kind;
                    t.dst = s.dst;
                    // S/N tiles entirely beyond rows_eval have nothing to do
                    if (s.kind == 1 && t.lo >= s.rows_eval) continue;
                } else {
                    t.kind = 0;
                    t.dst = other;
                }
                passes[p].push_back(t);
            }
        }
        if (!last) {
            const float* next_src = other;
            other = const_cast<float*>(src);
            src = next_src;
        }
    }
    return npass;
}

}  // namespace rt
