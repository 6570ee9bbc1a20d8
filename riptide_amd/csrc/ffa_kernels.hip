// FFA transform + boxcar S/N kernels for gfx950.
//
// One workgroup (16 wavefronts) computes one FfaTask: it stages the rows a
// tile of the merge tree depends on in LDS, applies the tree's merge levels
// in place (every level is read into registers, then written back after a
// barrier), and either writes the finished rows back to HBM or turns each
// finished row into per-width boxcar S/N without the row ever leaving LDS.
//
// Merge order and row map follow transforms.hpp:13-50: node of M rows =
// head (M>>1 rows) + tail; output row s adds head row h(s) and tail row t(s)
// rolled left by s - t(s), where h and t are computed in float32 as
// trunc(k*s + 0.5f) with k = (rows_child - 1)/(M - 1).  Every output element
// is one float32 add of the same two operands as the reference, so the
// transform is bit-identical.
#include <hip/hip_runtime.h>

#include "ffa_device.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "tasks.h"

namespace rt {

// LDS layout of the FFA task kernel (bytes are dynamic; see ffa_lds_bytes()):
//   float    zero[bins]                       all -0.0f (tail of a copied row)
//   float    data[cap_rows * bins]            staged rows
//   uint32_t tab[kFfaMaxLevels * cap_rows]    packed row map of every level
//   Seg      seg[2^(kMaxConeLevels+1)]        cone segments (cone tasks only)
struct Seg {
    int m;     // node rows (0: empty slot)
    int r0;    // node first row (transform-relative)
    int lo;    // first needed row (node-relative)
    int len;   // number of needed rows
    int base;  // first LDS row of this segment in its level's layout
};

__device__ __forceinline__ int seg_slot(int t, int j) { return (1 << t) - 1 + j; }

// Finds the segment slot of level t whose LDS range holds row p.
__device__ __forceinline__ int seg_find(const Seg* seg, int t, int p)
{
    const int first = (1 << t) - 1;
    int lo = 0, hi = (1 << t) - 1;
    // last slot with base <= p and len > 0; bases are non-decreasing
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[first + mid].base <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    // empty slots share the base of the next non-empty one; step back to the
    // slot that really contains p
    while (lo > 0 && (seg[first + lo].len == 0 || seg[first + lo].base + seg[first + lo].len <= p))
        --lo;
    return first + lo;
}

// Builds the cone of a tile top-down: level 0 is the tile itself; level t+1
// holds the rows of the children that level t reads.  Returns R_levels (rows
// at the bottom) in *bottom_rows.  Must be called by the whole workgroup.
__device__ void build_cone(Seg* seg, const FfaTask& tk, int* shared_int)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        Seg s;
        s.m = tk.node_m;
        s.r0 = tk.node_r0;
        s.lo = tk.lo;
        s.len = tk.hi - tk.lo;
        s.base = 0;
        seg[0] = s;
    }
    __syncthreads();
    for (int t = 0; t < tk.levels; ++t) {
        const int nslots = 1 << t;
        for (int j = tid; j < nslots; j += blockDim.x) {
            const Seg p = seg[seg_slot(t, j)];
            Seg a, b;
            a.m = 0; a.r0 = p.r0; a.lo = 0; a.len = 0; a.base = 0;
            b = a;
            if (p.len > 0 && p.m == 1) {
                a = p;  // single-row leaf: carried down unchanged
            } else if (p.len > 0 && p.m >= 2) {
                const int mh = p.m >> 1;
                const int mt = p.m - mh;
                const float kh = ((float)mh - 1.0f) / ((float)p.m - 1.0f);
                const float kt = ((float)mt - 1.0f) / ((float)p.m - 1.0f);
                const int s0 = p.lo, s1 = p.lo + p.len - 1;
                const int h0 = merge_row(kh, s0), h1 = merge_row(kh, s1);
                const int t0 = merge_row(kt, s0), t1 = merge_row(kt, s1);
                a.m = mh; a.r0 = p.r0; a.lo = h0; a.len = h1 - h0 + 1;
                b.m = mt; b.r0 = p.r0 + mh; b.lo = t0; b.len = t1 - t0 + 1;
            }
            seg[seg_slot(t + 1, 2 * j)] = a;
            seg[seg_slot(t + 1, 2 * j + 1)] = b;
        }
        __syncthreads();
        // exclusive scan of the children's lengths (<= 2^kMaxConeLevels slots)
        if (tid < kWave) {
            const int n = 2 << t;
            const int per = (n + kWave - 1) / kWave;
            const int first = seg_slot(t + 1, 0);
            int run = 0;
            for (int q = 0; q < per; ++q) {
                const int j = tid * per + q;
                if (j < n) run += seg[first + j].len;
            }
            int incl = run;
            for (int d = 1; d < kWave; d <<= 1) {
                const int o = __shfl_up(incl, d, kWave);
                if (tid >= d) incl += o;
            }
            int excl = incl - run;
            for (int q = 0; q < per; ++q) {
                const int j = tid * per + q;
                if (j < n) {
                    seg[first + j].base = excl;
                    excl += seg[first + j].len;
                }
            }
            if (tid == kWave - 1) *shared_int = incl;
        }
        __syncthreads();
    }
}

// Row map of one merge level of a whole-node (identity layout) task: row p of
// the node at relative depth t.
__device__ __forceinline__ void identity_row_map(int node_m, int t, int p, int bins, int* a,
                                                 int* b, int* sh)
{
    int r0 = 0, m = node_m;
    for (int d = 0; d < t && m > 1; ++d) {
        const int mh = m >> 1;
        if (p < r0 + mh) {
            m = mh;
        } else {
            r0 += mh;
            m -= mh;
        }
    }
    if (m <= 1) {
        *a = p * bins;
        *b = -1;
        *sh = 0;
        return;
    }
    const int mh = m >> 1;
    const int mt = m - mh;
    const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
    const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
    const int s = p - r0;
    const int h = merge_row(kh, s);
    const int tt = merge_row(kt, s);
    *a = (r0 + h) * bins;
    *b = (r0 + mh + tt) * bins;
    *sh = (s - tt) % bins;
}

// Row map of one merge level of a cone task (level t reads level t+1).
__device__ __forceinline__ void cone_row_map(const Seg* seg, int t, int p, int bins, int* a,
                                             int* b, int* sh)
{
    const int slot = seg_find(seg, t, p);
    const Seg g = seg[slot];
    const int j = slot - ((1 << t) - 1);
    const Seg ca = seg[seg_slot(t + 1, 2 * j)];
    const Seg cb = seg[seg_slot(t + 1, 2 * j + 1)];
    const int s = g.lo + (p - g.base);
    if (g.m == 1) {
        *a = (ca.base + (s - ca.lo)) * bins;
        *b = -1;
        *sh = 0;
        return;
    }
    const int mh = g.m >> 1;
    const int mt = g.m - mh;
    const float kh = ((float)mh - 1.0f) / ((float)g.m - 1.0f);
    const float kt = ((float)mt - 1.0f) / ((float)g.m - 1.0f);
    const int h = merge_row(kh, s);
    const int tt = merge_row(kt, s);
    *a = (ca.base + (h - ca.lo)) * bins;
    *b = (cb.base + (tt - cb.lo)) * bins;
    *sh = (s - tt) % bins;
}

// Packed per-row table entry of one merge level: head row (10 bits), tail
// row + 1 (10 bits, 0 = copy the head row) and circular shift (12 bits).
__device__ __forceinline__ uint32_t pack_row(int a, int b, int sh)
{
    return (uint32_t)a | ((uint32_t)(b + 1) << 10) | ((uint32_t)sh << 20);
}

// Copies `rows` rows of `bins` floats from global memory into LDS, one
// wavefront per row with the row's k column values loaded before they are
// stored (loads in flight instead of one round trip per value).
template <int NW>
__device__ __forceinline__ void load_rows(float* data, const float* src, const uint32_t* grow,
                                          int rows, int bins, int k, int wave, int lane)
{
    for (int p = wave; p < rows; p += NW) {
        const float* srow = src + (size_t)__builtin_amdgcn_readfirstlane(grow[p]) * bins;
        float* drow = data + p * bins;
        float v[kFfaLoadK];
#pragma unroll
        for (int q = 0; q < kFfaLoadK; ++q) {
            const int c = lane + q * kWave;
            if (q < k && c < bins) v[q] = srow[c];
        }
#pragma unroll
        for (int q = 0; q < kFfaLoadK; ++q) {
            const int c = lane + q * kWave;
            if (q < k && c < bins) drow[c] = v[q];
        }
        for (int c = lane + kFfaLoadK * kWave; c < bins; c += kWave) drow[c] = srow[c];
    }
}

// Rows in the layout of merge level t of a task (the rows that level writes).
__device__ __forceinline__ int level_rows(bool cone, const Seg* seg, int t, int rows)
{
    if (!cone) return rows;
    const int first = seg_slot(t, 0);
    const int n = 1 << t;
    return seg[first + n - 1].base + seg[first + n - 1].len;
}

// All merge levels of a task, bottom-up, in place, with at most K columns per
// lane (bins <= 64*K; exactly K when EXACT) known at compile time.
// Wavefront w owns rows w, w+NW, ...; lane l owns columns l + 64*q.  Per
// element: two LDS reads (head column c, tail column (c + shift) mod bins,
// the wrap chosen by one compare against a per-row scalar) and one add; the
// results stay in registers until the barrier, then overwrite the level in
// place.  Row -1 of `data` is the -0.0f row, the tail of copied rows.
// Kept out of line: one register allocation per column count.
template <int NW, int K, int STAGE, bool EXACT>
__device__ __forceinline__ int merge_levels_k(const Seg* seg, float* data,
                                                        const uint32_t* tab, int rows, int levels,
                                                        bool cone, int bins, int cap_rows,
                                                        int wave, int lane)
{
    constexpr int R = ffa_rows_per_wave(K, STAGE);
    static_assert(R >= 1, "column count beyond the register staging");
    const int kd = (bins + kWave - 1) / kWave;
    const bool last_ok = lane + (K - 1) * kWave < bins;
    for (int t = levels - 1; t >= 0; --t) {
        const int out_rows = level_rows(cone, seg, t, rows);
        const uint32_t* tl = tab + (size_t)t * cap_rows;
        const int nr = (out_rows - wave + NW - 1) / NW;
        // this wavefront's row-table entries, all read before the first use
        uint32_t ent[R];
#pragma unroll
        for (int i = 0; i < R; ++i) ent[i] = tl[min(wave + i * NW, out_rows - 1)];
        float v[R * K];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < nr) {
                const uint32_t e = __builtin_amdgcn_readfirstlane(ent[i]);
                const int a = (int)(e & 1023u);
                const int bt = (int)((e >> 10) & 1023u) - 1;
                const int sh = (int)(e >> 20);
                const float* hrow = data + a * bins + lane;
                // tail column c + sh before and after the wrap
                const float* trow = data + bt * bins + lane + sh;
                const float* twrap = trow - bins;
                const int lim = bins - sh;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    if (EXACT || q < kd) {
                        const float* tp = lane + q * kWave >= lim ? twrap : trow;
                        v[i * K + q] = hrow[q * kWave] + tp[q * kWave];
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < nr) {
                float* orow = data + (wave + i * NW) * bins + lane;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    if (EXACT) {
                        if (q < K - 1 || last_ok) orow[q * kWave] = v[i * K + q];
                    } else if (q < kd && lane + q * kWave < bins) {
                        orow[q * kWave] = v[i * K + q];
                    }
                }
            }
        }
        __syncthreads();
        rows = out_rows;
    }
    return rows;
}

// Dispatch on the columns per lane: exact instantiations up to 16 columns
// (bins <= 1024), then one per rows-per-wavefront class.
template <int NW, int STAGE>
__device__ __forceinline__ int merge_levels(const Seg* seg, float* data, const uint32_t* tab,
                                            int rows, int levels, bool cone, int bins,
                                            int cap_rows, int wave, int lane)
{
    const int k = (bins + kWave - 1) / kWave;
#define RT_MERGE_ARGS seg, data, tab, rows, levels, cone, bins, cap_rows, wave, lane
    switch (k) {
#define RT_MERGE_K(K)                                                                          \
    case K:                                                                                    \
        return merge_levels_k<NW, K, STAGE, true>(RT_MERGE_ARGS);
    RT_MERGE_K(1) RT_MERGE_K(2) RT_MERGE_K(3) RT_MERGE_K(4) RT_MERGE_K(5) RT_MERGE_K(6)
    RT_MERGE_K(7) RT_MERGE_K(8) RT_MERGE_K(9) RT_MERGE_K(10) RT_MERGE_K(11) RT_MERGE_K(12)
    RT_MERGE_K(13) RT_MERGE_K(14) RT_MERGE_K(15) RT_MERGE_K(16)
#undef RT_MERGE_K
    default: break;
    }
    // wide rows: R rows per wavefront of up to STAGE/R columns
    if (STAGE / 4 > 16 && k <= STAGE / 4)
        return merge_levels_k<NW, (STAGE / 4 > 16 ? STAGE / 4 : 16), STAGE, false>(RT_MERGE_ARGS);
    if (STAGE / 3 > 16 && k <= STAGE / 3)
        return merge_levels_k<NW, (STAGE / 3 > 16 ? STAGE / 3 : 16), STAGE, false>(RT_MERGE_ARGS);
    if (STAGE / 2 > 16 && k <= STAGE / 2)
        return merge_levels_k<NW, (STAGE / 2 > 16 ? STAGE / 2 : 16), STAGE, false>(RT_MERGE_ARGS);
    return merge_levels_k<NW, STAGE, STAGE, false>(RT_MERGE_ARGS);
#undef RT_MERGE_ARGS
}

// Runs one FfaTask with the whole workgroup; LDS contents are dead on return.
//
// The row tables of all merge levels are built first (one barrier).  Every
// level is then processed one row per wavefront: the row's packed table entry
// is one broadcast LDS read, lanes walk the columns c = lane + 64*q, and each
// lane keeps up to kStage results in registers until the barrier after which
// the level is written back in place.
template <int NT, int STAGE>
__device__ void run_ffa_task(const FfaTask& tk, float* lds, int* err)
{
    constexpr int NW = NT / kWave;
    const int bins = tk.bins;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int cap_rows = tk.pad;
    const int k = (bins + kWave - 1) / kWave;  // columns per lane

    // row -1 of data is all -0.0f; rows 0..cap_rows-1 hold the staged rows
    float* data = lds + bins;
    uint32_t* tab = reinterpret_cast<uint32_t*>(data + (size_t)cap_rows * bins);
    Seg* seg = reinterpret_cast<Seg*>(tab + (size_t)cap_rows * kFfaMaxLevels);
    __shared__ int s_bottom;

    for (int c = tid; c < bins; c += NT) lds[c] = -0.0f;

    // ---- bottom rows and the tables of every level
    int rows;  // rows in the current level's layout
    if (tk.cone) {
        build_cone(seg, tk, &s_bottom);
        rows = (tk.levels > 0) ? s_bottom : (tk.hi - tk.lo);
        if (rows > cap_rows || tk.levels > kFfaMaxLevels) {
            // the host sizes tiles so this cannot happen; never overrun LDS
            if (tid == 0) atomicOr(err, 1);
            __syncthreads();
            return;
        }
        // global source row of every bottom row (in the table slot of the
        // deepest level, which no merge level uses), then every level's map
        uint32_t* grow = tab + (size_t)(kFfaMaxLevels - 1) * cap_rows;
        const int L = tk.levels;
        for (int p = tid; p < rows; p += NT) {
            const Seg g = seg[seg_find(seg, L, p)];
            grow[p] = (uint32_t)(g.r0 + g.lo + (p - g.base));
        }
        for (int t = 0; t < L; ++t) {
            const int first = seg_slot(t, 0);
            const int n = 1 << t;
            const int rt = seg[first + n - 1].base + seg[first + n - 1].len;
            for (int p = tid; p < rt; p += NT) {
                int a, b, sh;
                cone_row_map(seg, t, p, bins, &a, &b, &sh);
                tab[(size_t)t * cap_rows + p] = pack_row(a / bins, b < 0 ? -1 : b / bins, sh);
            }
        }
        __syncthreads();
        load_rows<NW>(data, tk.src, grow, rows, bins, k, wave, lane);
    } else {
        rows = tk.node_m;
        if (tk.levels > kFfaMaxLevels) {
            if (tid == 0) atomicOr(err, 1);
            __syncthreads();
            return;
        }
        for (int t = 0; t < tk.levels; ++t)
            for (int p = tid; p < rows; p += NT) {
                int a, b, sh;
                identity_row_map(tk.node_m, t, p, bins, &a, &b, &sh);
                tab[(size_t)t * cap_rows + p] = pack_row(a / bins, b < 0 ? -1 : b / bins, sh);
            }
        // contiguous rows: a flat copy with several loads in flight per lane
        const float* src = tk.src + (size_t)tk.node_r0 * bins;
        const int total = rows * bins;
        for (int e0 = tid; e0 < total; e0 += NT * 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u * NT < total) v[u] = src[e0 + u * NT];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u * NT < total) data[e0 + u * NT] = v[u];
        }
    }
    __syncthreads();

    // ---- merge levels, bottom-up, in place through registers
    rows = merge_levels<NW, STAGE>(seg, data, tab, rows, tk.levels, tk.cone != 0, bins, cap_rows,
                                   wave, lane);

    // ---- outputs
    if (tk.kind == 0) {
        float* dst = tk.dst + ((size_t)tk.node_r0 + tk.lo) * bins;
        const int total = rows * bins;
        for (int e = tid; e < total; e += NT)
            dst[e] = data[e];
        __syncthreads();
        return;
    }

    // S/N of every finished row s = lo + p with s < rows_eval, in two
    // thread-level phases (no cross-lane traffic):
    //  1. one thread per row turns the row into its circular prefix sums in
    //     place, accumulating left to right in float-float (two-sum), so
    //     every partial sum is the float32 rounding of the exact sum, as the
    //     reference's float64 accumulator gives (kernels.hpp:73-90);
    //  2. one thread per (row, width) scans the row for the largest window
    //     sum and writes that width's S/N (snr.hpp:44-54).
    const int s_end = min(tk.hi, tk.rows_eval);
    const int nrows = max(0, s_end - tk.lo);
    for (int p = tid; p < nrows; p += NT) {
        float* row = data + (size_t)p * bins;
        float sh = 0.0f, sl = 0.0f;
        for (int i = 0; i < bins; ++i) {
            float s, e;
            two_sum(sh, row[i], s, e);
            sh = s;
            sl += e;
            const float r = sh + sl;  // rounded partial sum
            sl -= r - sh;             // keep (sh, sl) exact: sh + sl unchanged
            sh = r;
            row[i] = r;
        }
    }
    __syncthreads();
    const int nw = tk.nw;
    for (int job = tid; job < nrows * nw; job += NT) {
        const int p = job / nw;
        const int iw = job - p * nw;
        const float* cps = data + (size_t)p * bins;
        const int w = tk.widths[iw];
        const float total = cps[bins - 1];
        // i + w < bins: plain differences; beyond: the extension cps[j] + total
        float dm = -INFINITY;
        const int split = bins - w;
        for (int i = 0; i < split; ++i) dm = fmaxf(dm, cps[i + w] - cps[i]);
        for (int i = split; i < bins; ++i) dm = fmaxf(dm, (cps[i + w - bins] + total) - cps[i]);
        const float snr = __fmaf_rn(dm, tk.hb[2 * iw], -(tk.hb[2 * iw + 1] * total)) / tk.stdnoise;
        tk.snr[(size_t)(tk.lo + p) * nw + iw] = snr;
    }
    __syncthreads();
}

// One workgroup per task (the task list of one pass).
template <int NT, int STAGE, int WPS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void ffa_task_kernel(
    const FfaTask* __restrict__ tasks, int ntasks, int* err)
{
    extern __shared__ float lds[];
    if ((int)blockIdx.x >= ntasks) return;
    const FfaTask tk = tasks[blockIdx.x];
    run_ffa_task<NT, STAGE>(tk, lds, err);
}

// Node i of depth d of an m-row merge tree (row order); false when that slot
// lies under a leaf reached above depth d.
__device__ __forceinline__ bool tree_node(int m, int d, int i, int* r0, int* nm)
{
    int a = 0, M = m;
    for (int bit = d - 1; bit >= 0; --bit) {
        if (M <= 1) {
            if (i & ((1 << (bit + 1)) - 1)) return false;
            break;
        }
        const int mh = M >> 1;
        if ((i >> bit) & 1) {
            a += mh;
            M -= mh;
        } else {
            M = mh;
        }
    }
    *r0 = a;
    *nm = M;
    return true;
}

// Next work item of a transform for the persistent kernel, in pass order:
// pass 0 = whole nodes of depth depth[0] into buf_a, pass p >= 1 = cone tiles
// of the nodes of depth depth[p], alternating buf_a -> buf_b -> buf_a ...,
// the last pass producing S/N.  (p, i, lo) is the cursor.
__device__ __forceinline__ bool next_ffa_task(const FfaTransform& tr, const int32_t* depth,
                                              int& p, int& i, int& lo, float* buf_a,
                                              float* buf_b, FfaTask& tk)
{
    for (;;) {
        if (p >= tr.npass) return false;
        if (tr.npass == 1) {
            tk.src = tr.src;
            tk.node_r0 = 0;
            tk.node_m = tr.rows;
            tk.lo = 0;
            tk.hi = tr.rows;
            tk.levels = tr.levels0;
            tk.cone = 0;
            tk.kind = 1;
            p = 1;
            return true;
        }
        const int d = depth[p];
        if (i >= (1 << d)) {
            ++p;
            i = 0;
            lo = 0;
            continue;
        }
        int r0, nm;
        if (!tree_node(tr.rows, d, i, &r0, &nm)) {
            ++i;
            lo = 0;
            continue;
        }
        if (p == 0) {
            int h = 0;
            while ((1 << h) < nm) ++h;
            tk.src = tr.src;
            tk.dst = buf_a;
            tk.node_r0 = r0;
            tk.node_m = nm;
            tk.lo = 0;
            tk.hi = nm;
            tk.levels = h;
            tk.cone = 0;
            tk.kind = 0;
            ++i;
            return true;
        }
        const int L = depth[p - 1] - d;
        const int K = tr.cap - 2 * ((1 << L) - 1);
        const bool last = (p == tr.npass - 1);
        if (lo >= nm || (last && lo >= tr.rows_eval)) {
            ++i;
            lo = 0;
            continue;
        }
        tk.src = (p & 1) ? buf_a : buf_b;
        tk.dst = (p & 1) ? buf_b : buf_a;
        tk.node_r0 = r0;
        tk.node_m = nm;
        tk.lo = lo;
        tk.hi = min(lo + K, nm);
        tk.levels = L;
        tk.cone = 1;
        tk.kind = last ? 1 : 0;
        lo += K;
        return true;
    }
}

// Persistent batch kernel: every workgroup drains whole transforms from a
// device queue (largest first) and runs all of a transform's passes itself,
// its intermediate states in a private scratch slot, so no workgroup ever
// waits on another.  Rows written by one pass are read by the next only by
// this workgroup, ordered by the barrier that ends every task.
template <int NT, int STAGE, int WPS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void ffa_batch_kernel(
    const FfaTransform* __restrict__ trs, int ntr, int* __restrict__ queue, float* scratch,
    int64_t slot_floats, int* err)
{
    extern __shared__ float lds[];
    __shared__ int s_item;
    float* buf_a = scratch + (size_t)blockIdx.x * slot_floats;
    float* buf_b = buf_a + slot_floats / 2;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        // uniform across the workgroup: keep the descriptor in scalar registers
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= ntr) return;
        const FfaTransform tr = trs[item];
        FfaTask tk = {};
        tk.bins = tr.bins;
        tk.widths = tr.widths;
        tk.hb = tr.hb;
        tk.nw = tr.nw;
        tk.stdnoise = tr.stdnoise;
        tk.rows_eval = tr.rows_eval;
        tk.snr = tr.snr;
        tk.pad = tr.cap;
        int p = 0, i = 0, lo = 0;
        // the pass depths are indexed dynamically: read them from memory, so
        // that the descriptor itself stays in registers
        const int32_t* depth = trs[item].depth;
        while (next_ffa_task(tr, depth, p, i, lo, buf_a, buf_b, tk))
            run_ffa_task<NT, STAGE>(tk, lds, err);
    }
}

// Plain S/N of rows x cols profiles (python_bindings.cpp:121-160): one
// wavefront per row, the row staged in LDS.
__global__ __launch_bounds__(256) void snr_rows_kernel(const float* __restrict__ x, int rows,
                                                       int cols, const int32_t* widths,
                                                       const float* hb, int nw,
                                                       float stdnoise, float* __restrict__ out)
{
    extern __shared__ float lds[];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = blockDim.x / kWave;
    const int r = blockIdx.x * nwaves + wave;
    float* row = lds + (size_t)wave * cols;
    if (r < rows) {
        const float* src = x + (size_t)r * cols;
        for (int c = lane; c < cols; c += kWave)
            row[c] = src[c];
    }
    float own[kSnrMaxK];
    __syncthreads();
    if (r < rows)
        snr_prefix_inplace<kSnrMaxK>(row, cols, lane, own);
    __syncthreads();
    if (r < rows)
        snr_widths<kSnrMaxK>(row, cols, lane, own, widths, hb, nw, stdnoise, out + (size_t)r * nw);
}

// Same for rows longer than the register path supports: the prefix sums go
// to a global scratch row, one workgroup (one wavefront) per row.
__global__ __launch_bounds__(64) void snr_rows_long_kernel(const float* __restrict__ x, int rows,
                                                           int cols, const int32_t* widths,
                                                           const float* hb, int nw,
                                                           float stdnoise, float* scratch,
                                                           float* __restrict__ out)
{
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    if (r >= rows) return;
    float* cps = scratch + (size_t)r * cols;
    snr_prefix_long(x + (size_t)r * cols, cps, cols, lane);
    __syncthreads();
    snr_widths_long(cps, cols, lane, widths, hb, nw, stdnoise, out + (size_t)r * nw);
}

// One FFA merge level straight from HBM for blocks too wide for LDS: every
// output element adds the two operands the reference adds.  Launched once
// per tree depth, deepest first.
__global__ void ffa_level_global_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                        int64_t rows, int64_t cols, int depth)
{
    const int64_t total = rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = e / cols;
        const int64_t c = e - p * cols;
        // node at `depth` containing row p
        int64_t r0 = 0, m = rows;
        for (int d = 0; d < depth && m > 1; ++d) {
            const int64_t mh = m >> 1;
            if (p < r0 + mh) m = mh;
            else { r0 += mh; m -= mh; }
        }
        if (m <= 1) {
            dst[e] = src[e];
            continue;
        }
        const int64_t mh = m >> 1;
        const int64_t mt = m - mh;
        const float kh = ((float)mh - 1.0f) / ((float)m - 1.0f);
        const float kt = ((float)mt - 1.0f) / ((float)m - 1.0f);
        const int64_t s = p - r0;
        const int64_t h = merge_row(kh, (int)s);
        const int64_t t = merge_row(kt, (int)s);
        int64_t cc = c + (s - t) % cols;
        if (cc >= cols) cc -= cols;
        dst[e] = src[(r0 + h) * cols + c] + src[(r0 + mh + t) * cols + cc];
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

template <int NT, int STAGE, int WPS>
static hipError_t launch_tasks_shape(const FfaTask* d_tasks, int ntasks, int* d_err,
                                     hipStream_t stream, int lds)
{
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void*)ffa_task_kernel<NT, STAGE, WPS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((ffa_task_kernel<NT, STAGE, WPS>), dim3(ntasks), dim3(NT), lds, stream,
                       d_tasks, ntasks, d_err);
    return hipGetLastError();
}

hipError_t launch_ffa_tasks(const FfaTask* d_tasks, int ntasks, int* d_err, hipStream_t stream,
                            int shape)
{
    if (ntasks <= 0) return hipSuccess;
    const int lds = ffa_lds_bytes(shape);
    switch (shape) {
    case 0: return launch_tasks_shape<1024, 40, 4>(d_tasks, ntasks, d_err, stream, lds);
    case 1: return launch_tasks_shape<512, 40, 4>(d_tasks, ntasks, d_err, stream, lds);
    default: return launch_tasks_shape<512, 72, 2>(d_tasks, ntasks, d_err, stream, lds);
    }
}

template <int NT, int STAGE, int WPS>
static hipError_t launch_batch_shape(const FfaTransform* d_trs, int ntr, int* d_queue,
                                     float* d_scratch, int64_t slot_floats, int slots, int* d_err,
                                     hipStream_t stream, int lds)
{
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void*)ffa_batch_kernel<NT, STAGE, WPS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((ffa_batch_kernel<NT, STAGE, WPS>), dim3(slots), dim3(NT), lds, stream,
                       d_trs, ntr, d_queue, d_scratch, slot_floats, d_err);
    return hipGetLastError();
}

hipError_t launch_ffa_batch(const FfaTransform* d_trs, int ntr, int* d_queue, float* d_scratch,
                            int64_t slot_floats, int slots, int* d_err, hipStream_t stream,
                            int shape)
{
    if (ntr <= 0) return hipSuccess;
    const int lds = ffa_lds_bytes(shape);
    switch (shape) {
    case 0:
        return launch_batch_shape<1024, 40, 4>(d_trs, ntr, d_queue, d_scratch, slot_floats, slots,
                                               d_err, stream, lds);
    case 1:
        return launch_batch_shape<512, 40, 4>(d_trs, ntr, d_queue, d_scratch, slot_floats, slots,
                                              d_err, stream, lds);
    default:
        return launch_batch_shape<512, 72, 2>(d_trs, ntr, d_queue, d_scratch, slot_floats, slots,
                                              d_err, stream, lds);
    }
}

hipError_t launch_snr_rows(const float* x, int64_t rows, int cols, const int32_t* widths,
                           const float* hb, int nw, float stdnoise, float* out,
                           float* long_scratch, hipStream_t stream)
{
    if (rows <= 0) return hipSuccess;
    if (cols <= kWave * kSnrMaxK) {
        const int waves = 4;
        hipLaunchKernelGGL(snr_rows_kernel, dim3((unsigned)((rows + waves - 1) / waves)),
                           dim3(waves * kWave), (size_t)waves * cols * sizeof(float), stream, x,
                           (int)rows, cols, widths, hb, nw, stdnoise, out);
    } else {
        hipLaunchKernelGGL(snr_rows_long_kernel, dim3((unsigned)rows), dim3(kWave), 0, stream, x,
                           (int)rows, cols, widths, hb, nw, stdnoise, long_scratch, out);
    }
    return hipGetLastError();
}

hipError_t launch_ffa_level_global(const float* src, float* dst, int64_t rows, int64_t cols,
                                   int depth, hipStream_t stream)
{
    const int64_t total = rows * cols;
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(ffa_level_global_kernel, dim3((unsigned)g), dim3(256), 0, stream, src, dst,
                       rows, cols, depth);
    return hipGetLastError();
}

}  // namespace rt
