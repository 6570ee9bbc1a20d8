// FFA task kernels for gfx950 (device code and launch templates), compiled
// once per kernel variant by ffa_variant.hip.
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
#pragma once

#include <hip/hip_runtime.h>

#include "ffa_device.h"
#include "ffa_kernels.h"
#include "ffa_plan.h"
#include "task_iter.h"
#include "tasks.h"

namespace rt {

// Device-clock phase accounting of the FFA kernels, compiled in only with
// -DRT_PHASE_TIMING: thread 0 of a workgroup charges the cycles since its
// previous mark to phase `id` (0 queue and task iteration, 1 merge
// coefficients and cone, 2 staging and row tables, 3 / 4 merge levels of
// whole-node / cone tasks, 5 row store, 6 S/N prefix sums, 7 window maxima,
// 8 S/N writes); rt_phase_read adds up every workgroup's totals.
#ifdef RT_PHASE_TIMING
constexpr int kPhases = 16;
static __device__ unsigned long long g_rt_phase[kPhases];
__device__ __forceinline__ unsigned long long* rt_phase_lds()
{
    __shared__ unsigned long long s[kPhases + 1];
    return s;
}
#define RT_PHASE_START()                                                    \
    do {                                                                    \
        if (threadIdx.x == 0) {                                             \
            unsigned long long* s_ = rt_phase_lds();                        \
            for (int i_ = 0; i_ < kPhases; ++i_) s_[i_] = 0;                \
            s_[kPhases] = clock64();                                        \
        }                                                                   \
    } while (0)
#define RT_PHASE(id)                                                        \
    do {                                                                    \
        if (threadIdx.x == 0) {                                             \
            unsigned long long* s_ = rt_phase_lds();                        \
            const unsigned long long t_ = clock64();                        \
            s_[(id)] += t_ - s_[kPhases];                                   \
            s_[kPhases] = t_;                                               \
        }                                                                   \
    } while (0)
#define RT_PHASE_END()                                                      \
    do {                                                                    \
        if (threadIdx.x == 0) {                                             \
            unsigned long long* s_ = rt_phase_lds();                        \
            for (int i_ = 0; i_ < kPhases; ++i_) atomicAdd(&g_rt_phase[i_], s_[i_]); \
        }                                                                   \
    } while (0)
#else
#define RT_PHASE_START() do { } while (0)
#define RT_PHASE(id) do { } while (0)
#define RT_PHASE_END() do { } while (0)
#endif

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

// Merge coefficients k = (rows_child - 1)/(M - 1) (float32 division, as
// transforms.hpp:17-18) of every node size of a merge tree: the nodes at
// relative depth t of an M-row node have (M >> t) or (M >> t) + 1 rows, so
// two entries per depth; entry 2t + (m - (M >> t)).
struct MergeCoef {
    float kh[2 * (kFfaMaxLevels + 1)];
    float kt[2 * (kFfaMaxLevels + 1)];
};

__device__ __forceinline__ void fill_merge_coef(MergeCoef* mc, int M, int levels, int lane)
{
    if (lane < 2 * (levels + 1)) {
        const int t = lane >> 1;
        const int m = (M >> t) + (lane & 1);
        const int mh = m >> 1;
        const int mt = m - mh;
        mc->kh[lane] = m > 1 ? ((float)mh - 1.0f) / ((float)m - 1.0f) : 0.0f;
        mc->kt[lane] = m > 1 ? ((float)mt - 1.0f) / ((float)m - 1.0f) : 0.0f;
    }
}

__device__ __forceinline__ int coef_slot(int M, int t, int m) { return 2 * t + (m - (M >> t)); }

// Rows in the layout of merge level t of a cone (the rows that level
// writes; level `levels` is the staged bottom).
__device__ __forceinline__ int cone_level_rows(const Seg* seg, int t)
{
    const int first = seg_slot(t, 0);
    const int n = 1 << t;
    return seg[first + n - 1].base + seg[first + n - 1].len;
}

// Builds the cone of a tile top-down with one wavefront (no workgroup
// barrier): level 0 is the tile itself; level t+1 holds the rows of the
// children that level t reads.  Writes the rows of every level t <= levels
// to lrows[t].  A level has at most 2^kMaxConeLevels slots, two per lane.
__device__ __forceinline__ void build_cone_wave(Seg* seg, const FfaTask& tk, const MergeCoef* mc,
                                                int* lrows, int lane)
{
    static_assert((2 << (kMaxConeLevels - 1)) <= 2 * kWave, "cone level wider than a wavefront");
    const int M = tk.node_m;
    if (lane == 0) {
        Seg r;
        r.m = M;
        r.r0 = tk.node_r0;
        r.lo = tk.lo;
        r.len = tk.hi - tk.lo;
        r.base = 0;
        seg[0] = r;
    }
    for (int t = 0; t < tk.levels; ++t) {
        // the slots of level t were written by other lanes of this wavefront
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int j = lane;  // parent slot (2^t <= 64)
        Seg a, b;
        a.m = 0; a.r0 = 0; a.lo = 0; a.len = 0; a.base = 0;
        b = a;
        if (j < (1 << t)) {
            const Seg p = seg[seg_slot(t, j)];
            a.r0 = p.r0;
            b.r0 = p.r0;
            if (p.len > 0 && p.m == 1) {
                a = p;  // single-row leaf: carried down unchanged
            } else if (p.len > 0 && p.m >= 2) {
                const int mh = p.m >> 1;
                const int mt = p.m - mh;
                const int ks = coef_slot(M, t, p.m);
                const float kh = mc->kh[ks];
                const float kt = mc->kt[ks];
                const int s0 = p.lo, s1 = p.lo + p.len - 1;
                const int h0 = merge_row(kh, s0), h1 = merge_row(kh, s1);
                const int t0 = merge_row(kt, s0), t1 = merge_row(kt, s1);
                a.m = mh; a.r0 = p.r0; a.lo = h0; a.len = h1 - h0 + 1;
                b.m = mt; b.r0 = p.r0 + mh; b.lo = t0; b.len = t1 - t0 + 1;
            }
        }
        // exclusive scan of the children's lengths, children 2j and 2j+1 on lane j
        const int run = a.len + b.len;
        const int incl = wave_inclusive_sum(run);
        a.base = incl - run;
        b.base = a.base + a.len;
        if (j < (1 << t)) {
            seg[seg_slot(t + 1, 2 * j)] = a;
            seg[seg_slot(t + 1, 2 * j + 1)] = b;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane <= tk.levels) lrows[lane] = cone_level_rows(seg, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Packed per-row table entry of one merge level: head row (10 bits), tail
// row + 1 (10 bits, 0 = copy the head row) and circular shift (12 bits).
__device__ __forceinline__ uint32_t pack_row(int a, int b, int sh)
{
    return (uint32_t)a | ((uint32_t)(b + 1) << 10) | ((uint32_t)sh << 20);
}

// Row-table entries of every merge level for row p of a whole-node task (the
// node at relative depth t holding row p reads head row a and tail row b,
// rolled left by sh), walking the tree down from the root once.
__device__ __forceinline__ void identity_row_tables(int M, int levels, int p, int bins,
                                                    const MergeCoef* mc, uint32_t* tab,
                                                    int cap_rows)
{
    int r0 = 0, m = M;
    for (int t = 0; t < levels; ++t) {
        uint32_t e;
        if (m <= 1) {
            e = pack_row(p, -1, 0);
        } else {
            const int mh = m >> 1;
            const int ks = coef_slot(M, t, m);
            const int s = p - r0;
            const int h = merge_row(mc->kh[ks], s);
            const int tt = merge_row(mc->kt[ks], s);
            int sh = s - tt;
            while (sh >= bins) sh -= bins;
            e = pack_row(r0 + h, r0 + mh + tt, sh);
            if (s < mh) {
                m = mh;
            } else {
                r0 += mh;
                m -= mh;
            }
        }
        tab[(size_t)t * cap_rows + p] = e;
    }
}

// Row map of one merge level of a cone task (level t reads level t+1).
__device__ __forceinline__ void cone_row_map(const Seg* seg, int t, int p, int bins, int M,
                                             const MergeCoef* mc, int* a, int* b, int* sh)
{
    const int slot = seg_find(seg, t, p);
    const Seg g = seg[slot];
    const int j = slot - ((1 << t) - 1);
    const Seg ca = seg[seg_slot(t + 1, 2 * j)];
    const Seg cb = seg[seg_slot(t + 1, 2 * j + 1)];
    const int s = g.lo + (p - g.base);
    if (g.m == 1) {
        *a = ca.base + (s - ca.lo);
        *b = -1;
        *sh = 0;
        return;
    }
    const int ks = coef_slot(M, t, g.m);
    const int h = merge_row(mc->kh[ks], s);
    const int tt = merge_row(mc->kt[ks], s);
    *a = ca.base + (h - ca.lo);
    *b = cb.base + (tt - cb.lo);
    int r = s - tt;
    while (r >= bins) r -= bins;
    *sh = r;
}



// 32-bit LDS address of a pointer into LDS, and a load from one.
typedef __attribute__((address_space(3))) const float lds_cfloat;
__device__ __forceinline__ uint32_t lds_address(const float* p)
{
    return (uint32_t)(uintptr_t)(lds_cfloat*)p;
}
__device__ __forceinline__ float lds_load(uint32_t a) { return *(lds_cfloat*)(uintptr_t)a; }

// Everything of a task that depends on the row length, with at most K
// columns per lane (bins <= 64*K; exactly K when EXACT) known at compile time.
// LDS rows are `pitch` floats apart (odd, so that lanes reading the same
// column of consecutive rows hit distinct banks).  Wavefront w owns rows
// w, w+NW, ...; lane l owns columns l + 64*q.
template <int NW, int K, int STAGE, bool EXACT>
struct RowCore {
    static constexpr int R = ffa_rows_per_wave(K, STAGE);
    static_assert(R >= 1, "column count beyond the register staging");
    // rows per branch-free group of the merge loop (about 12 loads in flight)
    static constexpr int G = K >= 12 ? 1 : K >= 6 ? 2 : K >= 3 ? 4 : 8;

    // Column q of this lane exists in a row of `bins` columns.
    static __device__ __forceinline__ bool col_ok(int q, int kd, bool last_ok, int lane, int bins)
    {
        if (EXACT) return q < K - 1 || last_ok;
        return q < kd && lane + q * kWave < bins;
    }

    // Issues the global loads of this wavefront's rows (LDS rows wave + i*NW,
    // i < nr <= R) into registers: LDS row wave + i*NW is global row
    // readlane(grow_v, i) (cone tasks; lane i of this wavefront holds it) or
    // r0 + wave + i*NW (whole nodes).
    static __device__ __forceinline__ void load_issue(float (&v)[R * K], gcfloat* src, bool cone,
                                                      uint32_t grow_v, int r0, int nr, int bins,
                                                      int wave, int lane)
    {
        const int kd = (bins + kWave - 1) / kWave;
        const bool last_ok = lane + (K - 1) * kWave < bins;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < nr) {
                const int g = cone ? (int)__builtin_amdgcn_readlane(grow_v, i) : r0 + wave + i * NW;
                gcfloat* srow = src + (size_t)g * bins + lane;
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (col_ok(q, kd, last_ok, lane, bins)) v[i * K + q] = srow[q * kWave];
            }
        }
    }

    // Stores the registers of load_issue to the LDS rows.
    static __device__ __forceinline__ void load_commit(const float (&v)[R * K], float* data,
                                                       int nr, int bins, int pitch, int wave,
                                                       int lane)
    {
        const int kd = (bins + kWave - 1) / kWave;
        const bool last_ok = lane + (K - 1) * kWave < bins;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < nr) {
                float* drow = data + (wave + i * NW) * pitch + lane;
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (col_ok(q, kd, last_ok, lane, bins)) drow[q * kWave] = v[i * K + q];
            }
        }
    }

    // Writes LDS rows 0..rows-1 to consecutive global rows of dst.
    static __device__ __forceinline__ void store(gfloat* dst, const float* data, int rows,
                                                 int bins, int pitch, int wave, int lane)
    {
        const int kd = (bins + kWave - 1) / kWave;
        const bool last_ok = lane + (K - 1) * kWave < bins;
        for (int p = wave; p < rows; p += NW) {
            const float* srow = data + p * pitch + lane;
            gfloat* drow = dst + (size_t)p * bins + lane;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (col_ok(q, kd, last_ok, lane, bins)) drow[q * kWave] = srow[q * kWave];
        }
    }

    // All merge levels, bottom-up, in place.  Per element of a single level:
    // two LDS reads (head column c, tail column (c + shift) mod bins, the wrap
    // chosen by one compare against a per-row scalar) and one add; the
    // results stay in registers until the barrier, then overwrite the level
    // in place.  Rows of K >= 6 columns per lane take two levels per pass
    // (four reads and three adds per element, the row tables of both levels
    // composed per row): half the barriers and table rounds, 5/6 of the LDS
    // traffic.  Row -1 of `data` is the -0.0f row, the tail of copied rows.
    static __device__ __forceinline__ int merge(const int* lrows, float* data, const uint32_t* tab,
                                                int rows, int levels, bool cone, int bins,
                                                int pitch, int last_pitch, int cap_rows, int wave,
                                                int lane, gfloat* gdst)
    {
        const int kd = (bins + kWave - 1) / kWave;
        const bool last_ok = lane + (K - 1) * kWave < bins;
        for (int t = levels - 1; t >= 0;) {
            // levels t and t - 1 in one pass over LDS where two remain
            const bool two = K >= 6 && t >= 1;
            const int to = two ? t - 1 : t;  // the level this pass writes
            const int out_rows = cone ? lrows[to] : rows;
            const uint32_t* tl = tab + (size_t)to * cap_rows;
            const int nr = (out_rows - wave + NW - 1) / NW;
            // this wavefront's row-table entries, all read before the first use
            uint32_t ent[R];
#pragma unroll
            for (int i = 0; i < R; ++i) ent[i] = tl[min(wave + i * NW, out_rows - 1)];
            float v[R * K];
            if (two) {
                // output row p of level t-1 = X[a] + roll(X[b], sh) with X
                // = level t, X[r] = A[h(r)] + roll(A[t(r)], sh_r) and A =
                // level t+1 (in LDS), so
                //   out = (A[ha] + roll(A[ta], sha)) + (roll(A[hb], sh) + roll(A[tb], sh + shb))
                // in the reference's order of additions; a copied row (b =
                // -1) takes the -0.0 row for both terms of X[b]
                const uint32_t* tm = tab + (size_t)t * cap_rows;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    if (i < nr) {
                        const uint32_t e = __builtin_amdgcn_readfirstlane(ent[i]);
                        const int a = (int)(e & 1023u);
                        const int b = (int)((e >> 10) & 1023u) - 1;
                        const int sh = (int)(e >> 20);
                        const uint32_t fa = __builtin_amdgcn_readfirstlane(tm[a]);
                        const uint32_t fb = __builtin_amdgcn_readfirstlane(tm[b >= 0 ? b : 0]);
                        const int ha = (int)(fa & 1023u);
                        const int ta = (int)((fa >> 10) & 1023u) - 1;
                        const int sha = (int)(fa >> 20);
                        const int hb = b >= 0 ? (int)(fb & 1023u) : -1;
                        const int tb = b >= 0 ? (int)((fb >> 10) & 1023u) - 1 : -1;
                        const int shb = b >= 0 ? (int)(fb >> 20) : 0;
                        int shc = sh + shb;
                        if (shc >= bins) shc -= bins;
                        const float* hrow = data + ha * pitch + lane;
                        const uint32_t a1 = lds_address(data + ta * pitch + lane + sha);
                        const uint32_t b1 = lds_address(data + hb * pitch + lane + sh);
                        const uint32_t c1 = lds_address(data + tb * pitch + lane + shc);
                        uint32_t a2 = a1 - bins * 4, b2 = b1 - bins * 4, c2 = c1 - bins * 4;
                        asm volatile("" : "+v"(a2), "+v"(b2), "+v"(c2));
                        const int lima = bins - sha, limb = bins - sh, limc = bins - shc;
                        float x0[K], x1[K], x2[K], x3[K];
#pragma unroll
                        for (int q = 0; q < K; ++q) {
                            if (EXACT || q < kd) {
                                const int c = lane + q * kWave;
                                x0[q] = hrow[q * kWave];
                                x1[q] = lds_load((c >= lima ? a2 : a1) + q * kWave * 4);
                                x2[q] = lds_load((c >= limb ? b2 : b1) + q * kWave * 4);
                                x3[q] = lds_load((c >= limc ? c2 : c1) + q * kWave * 4);
                            }
                        }
                        asm volatile("" ::: "memory");
#pragma unroll
                        for (int q = 0; q < K; ++q)
                            if (EXACT || q < kd) v[i * K + q] = (x0[q] + x1[q]) + (x2[q] + x3[q]);
                    }
                }
            } else {
                // rows in groups of G without a branch between them, so that the
                // loads of a whole group are in flight together (a group may
                // run past this wavefront's last row: its entries are clamped to
                // a real row and its results are never written)
#pragma unroll
                for (int i0 = 0; i0 < R; i0 += G) {
                    if (i0 < nr) {
                        float hv[G * K], tv[G * K];
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const int i = i0 + g;
                            if (i >= R) break;
                            const uint32_t e = __builtin_amdgcn_readfirstlane(ent[i]);
                            const int a = (int)(e & 1023u);
                            const int bt = (int)((e >> 10) & 1023u) - 1;
                            const int sh = (int)(e >> 20);
                            const float* hrow = data + a * pitch + lane;
                            // LDS addresses of the tail column c + sh before and
                            // after the wrap; the second is hidden from the
                            // optimiser so that each element costs one compare
                            // and one select, the column step staying an
                            // immediate offset
                            const uint32_t t1 = lds_address(data + bt * pitch + lane + sh);
                            uint32_t t2 = t1 - bins * 4;
                            asm volatile("" : "+v"(t2));
                            const int lim = bins - sh;
#pragma unroll
                            for (int q = 0; q < K; ++q) {
                                if (EXACT || q < kd) {
                                    const uint32_t tb = lane + q * kWave >= lim ? t2 : t1;
                                    tv[g * K + q] = lds_load(tb + q * kWave * 4);
                                    hv[g * K + q] = hrow[q * kWave];
                                }
                            }
                        }
                        // every load of the group is issued before the first add
                        asm volatile("" ::: "memory");
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const int i = i0 + g;
                            if (i >= R) break;
#pragma unroll
                            for (int q = 0; q < K; ++q)
                                if (EXACT || q < kd) v[i * K + q] = hv[g * K + q] + tv[g * K + q];
                        }
                    }
                }
            }
            __syncthreads();
            if (K >= 6 && to == 0 && gdst != nullptr) {
                // the last level of a task whose rows go to HBM: written
                // straight from registers (row p to gdst + p*bins); no LDS
                // is read after the barrier above, so none is needed after.
                // Rows of fewer than 6 columns per lane keep the LDS store
                // phase: their many rows per wavefront would need more
                // address registers than the kernel has.
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    if (i < nr) {
                        gfloat* grow = gdst + (size_t)(wave + i * NW) * bins + lane;
#pragma unroll
                        for (int q = 0; q < K; ++q)
                            if (col_ok(q, kd, last_ok, lane, bins)) grow[q * kWave] = v[i * K + q];
                    }
                }
                return out_rows;
            }
            // the last level may be laid out at another pitch (rows with
            // room for their S/N prefix extension)
            const int opitch = to == 0 ? last_pitch : pitch;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                if (i < nr) {
                    float* orow = data + (wave + i * NW) * opitch + lane;
#pragma unroll
                    for (int q = 0; q < K; ++q)
                        if (col_ok(q, kd, last_ok, lane, bins)) orow[q * kWave] = v[i * K + q];
                }
            }
            __syncthreads();
            rows = out_rows;
            t = to - 1;
        }
        return rows;
    }

    // Staging, merge levels and (kind 0) store of one task; returns the rows
    // of the final level, or -1 if the rows exceed the register staging
    // (tasks are sized so that they do not).  The row tables are built by
    // `tables` while this wavefront's row loads are in flight.
    template <class Tables, class Setup>
    static __device__ __forceinline__ int run(const FfaTask& tk, const int* lrows, float* data,
                                              const uint32_t* tab, uint32_t grow_v, int rows,
                                              int pitch, int last_pitch, int wave, int lane,
                                              Tables&& tables, Setup&& setup)
    {
        const int bins = __builtin_amdgcn_readfirstlane(tk.bins);
        if (rows > R * NW) {
            setup();
            return -1;
        }
        const int nr = (rows - wave + NW - 1) / NW;
        float v[R * K];
        load_issue(v, as_global(tk.src), tk.cone != 0, grow_v,
                   __builtin_amdgcn_readfirstlane(tk.node_r0), nr, bins, wave, lane);
        // the merge coefficients of a whole-node task (one wavefront, one
        // barrier) while its loads are in flight; a cone task needed them
        // before its loads
        setup();
        tables();
        load_commit(v, data, nr, bins, pitch, wave, lane);
        __syncthreads();
        RT_PHASE(2);
        const int levels = __builtin_amdgcn_readfirstlane(tk.levels);
        gfloat* dst = tk.kind == 0 ? as_global(tk.dst) + ((size_t)tk.node_r0 + tk.lo) * bins : nullptr;
        rows = merge(lrows, data, tab, rows, levels, tk.cone != 0, bins, pitch, last_pitch,
                     __builtin_amdgcn_readfirstlane(tk.pad), wave, lane, dst);
        RT_PHASE(tk.cone ? 4 : 3);
        if (tk.kind == 0 && (levels == 0 || K < 6)) {
            // no merge level (the staged rows themselves), or the last level
            // was written back to LDS
            store(dst, data, rows, bins, pitch, wave, lane);
            __syncthreads();
            RT_PHASE(5);
        }
        return rows;
    }
};

// Dispatch on the columns per lane k = ceil(bins/64) within kernel variant V
// (ffa_variant()): variant V < kFfaWideVariant holds the exact code of
// k = 2V and 2V+1; the wide variant one instantiation per rows-per-wavefront
// class for k > kFfaMaxExactK.  Returns -1 for a k the variant does not hold
// (the host groups work by variant, so this does not happen).
template <int NW, int STAGE, int V, class Tables, class Setup>
__device__ __forceinline__ int run_rows(const FfaTask& tk, const int* lrows, float* data,
                                        const uint32_t* tab, uint32_t grow_v, int rows,
                                        int pitch, int last_pitch, int wave, int lane,
                                        Tables&& tables, Setup&& setup)
{
    const int k = (tk.bins + kWave - 1) / kWave;
#define RT_CORE_ARGS tk, lrows, data, tab, grow_v, rows, pitch, last_pitch, wave, lane, tables, setup
    if constexpr (V < kFfaWideVariant) {
        constexpr int KA = 2 * V, KB = 2 * V + 1;
        static_assert(KB <= kFfaMaxExactK, "variant beyond the exact column counts");
        if constexpr (KA >= 1) {
            if (k == KA) return RowCore<NW, KA, STAGE, true>::run(RT_CORE_ARGS);
        }
        if (k == KB) return RowCore<NW, KB, STAGE, true>::run(RT_CORE_ARGS);
        return -1;
    } else {
        constexpr int X = kFfaMaxExactK;
        constexpr int K4 = STAGE / 4 > X ? STAGE / 4 : X + 1;
        constexpr int K3 = STAGE / 3 > X ? STAGE / 3 : X + 1;
        constexpr int K2 = STAGE / 2 > X ? STAGE / 2 : X + 1;
        if (k <= X || k > STAGE) return -1;
        if (STAGE / 4 > X && k <= STAGE / 4) return RowCore<NW, K4, STAGE, false>::run(RT_CORE_ARGS);
        if (STAGE / 3 > X && k <= STAGE / 3) return RowCore<NW, K3, STAGE, false>::run(RT_CORE_ARGS);
        if (STAGE / 2 > X && k <= STAGE / 2) return RowCore<NW, K2, STAGE, false>::run(RT_CORE_ARGS);
        return RowCore<NW, STAGE, STAGE, false>::run(RT_CORE_ARGS);
    }
#undef RT_CORE_ARGS
}

// Order-preserving map of a float to an int (and back: the map is its own
// inverse), so that LDS integer max gives the float max.
__device__ __forceinline__ int ordered_int(float f)
{
    const int i = __float_as_int(f);
    return i ^ ((i >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float ordered_float(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF)); }

// Window maxima over window starts [i0, i1) of one row for NWC widths, all
// windows inside the row (i1 + wmax <= bins): each prefix value is read once
// for all widths.  Every lane of a wavefront works on the same columns of a
// different row, so all branches are uniform.
template <int NWC>
__device__ __forceinline__ void window_max_chunk(const float* cps, int i0, int i1,
                                                 const int (&w)[kSnrMaxWidths],
                                                 float (&m)[kSnrMaxWidths])
{
    const int plain_end = i1;
    int i = i0;
    for (; i + 4 <= plain_end; i += 4) {
        float c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = cps[i + u];
        // widths four at a time: all loads of a group issued before the
        // first use (the empty asm keeps the compiler from interleaving
        // each load with its use, which would wait on every load)
#pragma unroll
        for (int j0 = 0; j0 < NWC; j0 += 4) {
            float q[4][4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
                if (j0 + jj < NWC) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[jj][u] = cps[i + u + w[j0 + jj]];
                }
            asm volatile("" ::: "memory");
            // two three-operand maxima per four windows
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
                if (j0 + jj < NWC) {
                    const int j = j0 + jj;
                    m[j] = fmaxf(fmaxf(m[j], q[jj][0] - c[0]), q[jj][1] - c[1]);
                    m[j] = fmaxf(fmaxf(m[j], q[jj][2] - c[2]), q[jj][3] - c[3]);
                }
        }
    }
    for (; i < plain_end; ++i) {
        const float c0 = cps[i];
#pragma unroll
        for (int j = 0; j < NWC; ++j) m[j] = fmaxf(m[j], cps[i + w[j]] - c0);
    }
}

// Window maxima of the first NWT default widths (kDefaultWidths) over the
// starts [i0, i1): the same maxima as window_max_chunk, with compile-time
// offsets.  The small widths (1, 2, 3, 4, 6, 9, 13, 19) of four consecutive
// starts read only prefix values i .. i + 22: that block serves the start
// values and all of them, and it slides by four starts per step, so only its
// last four values are loaded; the larger widths load their own four values.
// Per four starts 4 + 4 (NWT - 8) loads instead of 4 (NWT + 1).
template <int NWT>
__device__ __forceinline__ void window_max_default(const float* cps, int i0, int i1,
                                                   float (&m)[kSnrMaxWidths])
{
    constexpr int NS = NWT < 8 ? NWT : 8;  // widths served by the block
    constexpr int NB = kDefaultWidths[NS - 1] + 4;  // block cps[i .. i + NB)
    constexpr int NL = NWT - NS;  // widths with their own loads
    int i = i0;
    float blk[NB];
    if (i + 4 <= i1) {
#pragma unroll
        for (int k = 0; k < NB - 4; ++k) blk[k + 4] = cps[i + k];
    }
    for (; i + 4 <= i1; i += 4) {
        // slide the block by four starts: NB - 4 values carried over
#pragma unroll
        for (int k = 0; k < NB - 4; ++k) blk[k] = blk[k + 4];
#pragma unroll
        for (int k = NB - 4; k < NB; ++k) blk[k] = cps[i + k];
        float q[NL > 0 ? NL : 1][4];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
#pragma unroll
            for (int u = 0; u < 4; ++u) q[j][u] = cps[i + u + kDefaultWidths[NS + j]];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int W = kDefaultWidths[j];
            m[j] = fmaxf(fmaxf(m[j], blk[W] - blk[0]), blk[W + 1] - blk[1]);
            m[j] = fmaxf(fmaxf(m[j], blk[W + 2] - blk[2]), blk[W + 3] - blk[3]);
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            m[NS + j] = fmaxf(fmaxf(m[NS + j], q[j][0] - blk[0]), q[j][1] - blk[1]);
            m[NS + j] = fmaxf(fmaxf(m[NS + j], q[j][2] - blk[2]), q[j][3] - blk[3]);
        }
    }
    for (; i < i1; ++i) {
        const float c0 = cps[i];
#pragma unroll
        for (int j = 0; j < NWT; ++j) m[j] = fmaxf(m[j], cps[i + kDefaultWidths[j]] - c0);
    }
}

// Per-width boxcar S/N of `nrows` finished rows staged in LDS (row p at
// data + p*pitch), written to tk.snr.  `scratch` is free LDS.
//  1. circular prefix sums in place (kernels.hpp:73-90): each row is cut
//     into C chunks; one thread per (chunk, row), chunk-major so that the
//     lanes of a wavefront take the same chunk of consecutive rows, with C
//     chosen so that about every thread of the workgroup has an item.
//     Chunk totals in float64 go to LDS, then every item rescans its chunk
//     from the float64 sum of the chunks before it (which it adds up itself,
//     in chunk order), writing each partial sum rounded to float32, as the
//     reference's float64 accumulator does.
//  2. window maxima: the same (chunk, row) threads, all widths per thread
//     (each prefix value read once per width), combined across chunks with
//     LDS integer max; then one thread per (row, width) writes the S/N.
template <int NT>
__device__ __forceinline__ void snr_rows(const FfaTask& tk, float* data, int pitch, int nrows,
                                         float* scratch, int scratch_floats, int* s_w,
                                         float* s_hb, int ext_w)
{
    constexpr int NW = NT / kWave;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int bins = tk.bins;
    const int nw = tk.nw;
    if (nrows <= 0) return;
    // (chunk, row) work items, item it = c * nrows + p: every thread takes
    // one item where C * nrows <= NT, so no lane idles for want of a row.
    // Chunk c covers [chunk_lo(c), chunk_lo(c + 1)) in whole groups of 4
    // columns (the last chunk also takes the remainder); the lanes of one
    // wavefront work on consecutive rows of one or two chunks.
    int C = max(1, NT / nrows);
    C = min(C, max(1, bins / 8));  // chunks of >= 8 columns
    // float64 chunk totals: [C][nrows] in the scratch, or (rows with room
    // for the prefix extension, ext_w > 0) the C totals of row p in that
    // row's extension, which the extension overwrites once they are used
    if (ext_w)
        C = min(C, max(1, (ext_w - 1) / 2));
    else
        C = min(C, max(1, (scratch_floats - 1) / (2 * nrows)));
    const int nitems = C * nrows;
    double* tot = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(scratch) + 7) &
                                            ~static_cast<uintptr_t>(7));
    auto tot_at = [&](int c, int p) -> double* {
        if (ext_w) {
            const uintptr_t e = reinterpret_cast<uintptr_t>(data + (size_t)p * pitch + bins);
            return reinterpret_cast<double*>((e + 7) & ~static_cast<uintptr_t>(7)) + c;
        }
        return tot + (size_t)c * nrows + p;
    };
    // widths and (h + b, b) pairs of the transform: loaded now, stored to
    // LDS after the first prefix pass, so that their latency is hidden
    const int nwl = min(nw, kSnrMaxWidths);
    int w_ld = 0;
    float hb_ld = 0.0f;
    if (wave == NW - 1) {
        if (lane < nwl) w_ld = as_global(tk.widths)[lane];
        if (lane < 2 * nwl) hb_ld = as_global(tk.hb)[lane];
    }
    const int g4 = bins >> 2;
    // a thread's item (chunk c, row p) and its chunk's bounds: worked out
    // once for every pass where each thread has at most one item (integer
    // divisions are long instruction sequences)
    const bool one = nitems <= NT;
    const int c0 = tid / nrows;
    const int p0 = tid - c0 * nrows;
    const int j00 = 4 * ((c0 * g4) / C);
    const int j01 = c0 == C - 1 ? bins : 4 * (((c0 + 1) * g4) / C);
    auto item = [&](int it, int& c, int& p) {
        if (one) {
            c = c0;
            p = p0;
        } else {
            c = it / nrows;
            p = it - c * nrows;
        }
    };
    auto bounds = [&](int c, int& j, int& je) {
        if (one) {
            j = j00;
            je = j01;
        } else {
            j = 4 * ((c * g4) / C);
            je = c == C - 1 ? bins : 4 * (((c + 1) * g4) / C);
        }
    };
    // (1) chunk totals, accumulated in float64 as the reference does
    for (int it = tid; it < nitems; it += NT) {
        int c, p, j, je;
        item(it, c, p);
        bounds(c, j, je);
        const float* row = data + (size_t)p * pitch;
        double acc = 0.0;
        for (; j + 4 <= je; j += 4) {
            float x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = row[j + u];
            asm volatile("" ::: "memory");
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += (double)x[u];
        }
        for (; j < je; ++j) acc += (double)row[j];
        *tot_at(c, p) = acc;
    }
    if (wave == NW - 1) {
        if (lane < nwl) s_w[lane] = w_ld;
        if (lane < 2 * nwl) s_hb[lane] = hb_ld;
    }
    __syncthreads();
    // (2) every chunk rescanned in float64 from the sum of the chunks before
    // it (added in chunk order by the item itself), each partial sum rounded
    // to float32 (kernels.hpp:75-83).  The chunk totals stay in place: items
    // only read them, so no barrier is needed before the rows are
    // overwritten.
    for (int it = tid; it < nitems; it += NT) {
        int c, p, j, je;
        item(it, c, p);
        bounds(c, j, je);
        double acc = 0.0;
        for (int q = 0; q < c; ++q) acc += *tot_at(q, p);
        float* row = data + (size_t)p * pitch;
        for (; j + 4 <= je; j += 4) {
            float x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = row[j + u];
            asm volatile("" ::: "memory");
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc += (double)x[u];
                x[u] = (float)acc;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) row[j + u] = x[u];
        }
        for (; j < je; ++j) {
            acc += (double)row[j];
            row[j] = (float)acc;
        }
    }
    __syncthreads();
    RT_PHASE(6);

    gcfloat* hb = as_global(tk.hb);
    gcint* gw = as_global(tk.widths);
    int* mx = reinterpret_cast<int*>(scratch);  // [nrows][nw] window maxima
    if (nw >= 1 && nw <= kSnrMaxWidths && nrows * nw <= scratch_floats) {
        for (int e = tid; e < nrows * nw; e += NT) mx[e] = ordered_int(-INFINITY);
        // widths and (h + b, b) pairs were stored to LDS after the first
        // prefix pass; one LDS read per lane (lane j: width j), then the
        // widths from the lanes, so that the reads are not serialised
        const int w_lane = s_w[min(lane, nw - 1)];
        int w[kSnrMaxWidths];
#pragma unroll
        for (int j = 0; j < kSnrMaxWidths; ++j) w[j] = __builtin_amdgcn_readlane(w_lane, j);
        // the widest window (the widths need not be in order)
        int wmax = w[0];
#pragma unroll
        for (int j = 1; j < kSnrMaxWidths; ++j) wmax = max(wmax, w[j]);
        // default width trials (7 or more): the compile-time path
        bool dflt = nw >= 7;
#pragma unroll
        for (int j = 0; j < kSnrMaxWidths; ++j)
            if (j < nw && w[j] != kDefaultWidths[j]) dflt = false;
        if (ext_w) {
            // the circular extension of the prefix sums, cps[bins + j] =
            // cps[j] + total for j < wmax (kernels.hpp:86-101, float32), over
            // the chunk totals, which are no longer read
            for (int e = tid; e < nrows * ext_w; e += NT) {
                const int p = e / ext_w;
                const int j = e - p * ext_w;
                float* row = data + (size_t)p * pitch;
                row[bins + j] = row[j] + row[bins - 1];
            }
        }
        __syncthreads();
        // (a) window starts [0, P) whose windows all stay inside the row (or
        // its extension), cut into the same C chunks of whole groups of 4,
        // all widths per thread: every start when the rows carry the
        // extension, else [0, bins - wmax)
        const int P = ext_w ? bins : max(0, bins - wmax);
        const int q4 = P >> 2;
        const int i00 = 4 * ((c0 * q4) / C);
        const int i01 = c0 == C - 1 ? P : 4 * (((c0 + 1) * q4) / C);
        for (int it = tid; it < nitems; it += NT) {
            int c, p;
            item(it, c, p);
            const float* cps = data + (size_t)p * pitch;
            const int i0 = one ? i00 : 4 * ((c * q4) / C);
            const int i1 = one ? i01 : c == C - 1 ? P : 4 * (((c + 1) * q4) / C);
            if (i0 >= i1) continue;
            float m[kSnrMaxWidths];
#pragma unroll
            for (int j = 0; j < kSnrMaxWidths; ++j) m[j] = -INFINITY;
            if (dflt) {
                switch (nw) {
#define RT_WIND(N) \
    case N: window_max_default<N>(cps, i0, i1, m); break;
                    RT_WIND(7) RT_WIND(8) RT_WIND(9) RT_WIND(10) RT_WIND(11) RT_WIND(12)
                    RT_WIND(13) RT_WIND(14) RT_WIND(15) RT_WIND(16)
#undef RT_WIND
                    default: break;
                }
            } else switch (nw) {
#define RT_WIN(N) \
    case N: window_max_chunk<N>(cps, i0, i1, w, m); break;
                RT_WIN(1) RT_WIN(2) RT_WIN(3) RT_WIN(4) RT_WIN(5) RT_WIN(6) RT_WIN(7) RT_WIN(8)
                RT_WIN(9) RT_WIN(10) RT_WIN(11) RT_WIN(12) RT_WIN(13) RT_WIN(14) RT_WIN(15)
                RT_WIN(16)
#undef RT_WIN
                default: break;
            }
#pragma unroll
            for (int j = 0; j < kSnrMaxWidths; ++j)
                if (j < nw) atomicMax(&mx[p * nw + j], ordered_int(m[j]));
        }
        // (b) without the extension, the last wmax window starts [P, bins),
        // one thread per (row, width), width-major.  The window at k ends inside the row while k <
        // split = bins - wj, else on the extension cps[j] + total (kernels.hpp:
        // 88-100).  Eight starts per step with all sixteen loads issued
        // before the first use; the step that crosses split and the last
        // partial step take the per-start form (starts past the row clamped
        // to its last start, which leaves the maximum unchanged).
        for (int job = tid; job < (ext_w ? 0 : nrows * nw); job += NT) {
            const int iw = job / nrows;
            const int p = job - iw * nrows;
            const float* cps = data + (size_t)p * pitch;
            const int wj = s_w[iw];
            const float total = cps[bins - 1];
            const int split = bins - wj;
            float d = -INFINITY;
            int k = P;
            auto mixed = [&](int k0) {
                float a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int st = min(k0 + u, bins - 1);
                    const int e = st + wj;
                    a[u] = cps[st];
                    b[u] = cps[e < bins ? e : e - bins];
                }
                asm volatile("" ::: "memory");
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int st = min(k0 + u, bins - 1);
                    const float v = st < split ? b[u] : b[u] + total;
                    d = fmaxf(d, v - a[u]);
                }
            };
            for (; k + 8 <= split; k += 8) {
                float a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = cps[k + u];
                    b[u] = cps[k + u + wj];
                }
                asm volatile("" ::: "memory");
                d = fmaxf(d, fmaxf(fmaxf(fmaxf(b[0] - a[0], b[1] - a[1]), fmaxf(b[2] - a[2], b[3] - a[3])),
                                   fmaxf(fmaxf(b[4] - a[4], b[5] - a[5]), fmaxf(b[6] - a[6], b[7] - a[7]))));
            }
            if (k < split) {
                mixed(k);
                k += 8;
            }
            for (; k + 8 <= bins; k += 8) {
                float a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = cps[k + u];
                    b[u] = cps[k + u - split];
                }
                asm volatile("" ::: "memory");
#pragma unroll
                for (int u = 0; u < 8; ++u) d = fmaxf(d, (b[u] + total) - a[u]);
            }
            if (k < bins) mixed(k);
            atomicMax(&mx[p * nw + iw], ordered_int(d));
        }
        __syncthreads();
        RT_PHASE(7);
        for (int e = tid; e < nrows * nw; e += NT) {
            const int p = e / nw;
            const int iw = e - p * nw;
            const float total = data[(size_t)p * pitch + bins - 1];
            const float dm = ordered_float(mx[e]);
            const float snr =
                __fmaf_rn(dm, s_hb[2 * iw], -(s_hb[2 * iw + 1] * total)) / tk.stdnoise;
            as_global(tk.snr)[(size_t)(tk.lo + p) * nw + iw] = snr;
        }
        __syncthreads();
        RT_PHASE(8);
        return;
    }
    // many widths: one thread per (row, width), width-major (uniform loop
    // bounds; the odd row pitch keeps the lanes' LDS reads on distinct banks)
    for (int job = tid; job < nrows * nw; job += NT) {
        const int iw = job / nrows;
        const int p = job - iw * nrows;
        const float* cps = data + (size_t)p * pitch;
        const int w = gw[iw];
        const float total = cps[bins - 1];
        const int split = bins - w;
        float d0 = -INFINITY, d1 = -INFINITY;
        int i = 0;
        for (; i + 2 <= split; i += 2) {
            d0 = fmaxf(d0, cps[i + w] - cps[i]);
            d1 = fmaxf(d1, cps[i + 1 + w] - cps[i + 1]);
        }
        for (; i < split; ++i) d0 = fmaxf(d0, cps[i + w] - cps[i]);
        // windows crossing the end: the extension cps[j] + total
        for (; i < bins; ++i) d1 = fmaxf(d1, (cps[i - split] + total) - cps[i]);
        const float dm = fmaxf(d0, d1);
        const float snr = __fmaf_rn(dm, hb[2 * iw], -(hb[2 * iw + 1] * total)) / tk.stdnoise;
        as_global(tk.snr)[(size_t)(tk.lo + p) * nw + iw] = snr;
    }
    __syncthreads();
}

// Runs one FfaTask with the whole workgroup; LDS contents are dead on return.
//
// Wavefront 0 computes the merge coefficients and the cone (one barrier);
// every wavefront then issues the global loads of its rows, the row tables of
// all merge levels are built while they are in flight, the rows are stored to
// LDS (one barrier), merged level by level and either written back to HBM or
// turned into per-width boxcar S/N without leaving LDS.
template <int NT, int STAGE, int V>
__device__ __forceinline__ void run_ffa_task(const FfaTask& tk, float* lds, int* err)
{
    constexpr int NW = NT / kWave;
    // task fields are uniform: keep them in scalar registers
    const int bins = __builtin_amdgcn_readfirstlane(tk.bins);
    const int pitch = ffa_pitch(bins);
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int cap_rows = __builtin_amdgcn_readfirstlane(tk.pad);

    // row -1 of data is all -0.0f; rows 0..cap_rows-1 hold the staged rows
    float* data = lds + pitch;
    uint32_t* tab = reinterpret_cast<uint32_t*>(data + (size_t)cap_rows * pitch);
    Seg* seg = reinterpret_cast<Seg*>(tab + (size_t)cap_rows * kFfaMaxLevels);
    // rows of every merge level of a cone task
    __shared__ int s_lrows[kMaxConeLevels + 1];

    __shared__ MergeCoef s_coef;
    const int levels = __builtin_amdgcn_readfirstlane(tk.levels);
    const int M = __builtin_amdgcn_readfirstlane(tk.node_m);
    if (levels > kFfaMaxLevels || (tk.cone && levels > kMaxConeLevels) ||
        (!tk.cone && M > cap_rows)) {
        // the host sizes tasks so this cannot happen; never overrun LDS
        if (tid == 0) atomicOr(err, 1);
        return;
    }
    // the task's tile record: its cone and row tables come from memory
    const int32_t* rec = tk.rec;

    for (int c = tid; c < bins; c += NT) lds[c] = -0.0f;

    __shared__ int s_w[kSnrMaxWidths];
    __shared__ float s_hb[2 * kSnrMaxWidths];

    // ---- merge coefficients and the cone (one wavefront, one barrier):
    // before the row loads of a cone task (they need the cone), while they
    // are in flight for a whole-node task.  With a record: only the level
    // rows, read before the staging barrier that precedes their first use.
    auto setup = [&]() {
        if (rec) {
            if (tk.cone && wave == 0 && lane <= levels) s_lrows[lane] = as_global(rec)[1 + lane];
            return;
        }
        if (wave == 0) {
            fill_merge_coef(&s_coef, M, levels, lane);
            if (tk.cone) {
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                build_cone_wave(seg, tk, &s_coef, s_lrows, lane);
            }
        }
        __syncthreads();
    };
    if (tk.cone) setup();
    RT_PHASE(1);
    int rows;  // rows in the current level's layout
    uint32_t grow_v = 0;
    if (tk.cone) {
        rows = (levels == 0) ? (tk.hi - tk.lo)
                             : __builtin_amdgcn_readfirstlane(rec ? as_global(rec)[0] : s_lrows[levels]);
        if (rows > cap_rows) {
            if (tid == 0) atomicOr(err, 1);
            __syncthreads();
            return;
        }
        // lane i holds the global source row of this wavefront's i-th LDS
        // row (wave + i*NW), the row its loads are issued from
        const int p = wave + lane * NW;
        if (p < rows) {
            if (rec) {
                grow_v = (uint32_t)(tk.node_r0 + as_global(rec)[kRecHeader + p]);
            } else {
                const Seg g = seg[seg_find(seg, levels, p)];
                grow_v = (uint32_t)(g.r0 + g.lo + (p - g.base));
            }
        }
    } else {
        rows = M;
    }
    // the row tables of every level, built (or copied from the record) while
    // the row loads are in flight
    const int rows0 = rows;
    auto tables = [&, rows0]() {
        if (rec) {
            gcint* src = as_global(rec) + kRecHeader + cap_rows;
            for (int e = tid; e < levels * cap_rows; e += NT) tab[e] = (uint32_t)src[e];
        } else if (tk.cone) {
            // one flat loop over (level, row) so that every thread has work
            for (int e = tid; e < levels * cap_rows; e += NT) {
                const int t = e / cap_rows;
                const int p = e - t * cap_rows;
                if (p < s_lrows[t]) {
                    int a, b, sh;
                    cone_row_map(seg, t, p, bins, M, &s_coef, &a, &b, &sh);
                    tab[e] = pack_row(a, b, sh);
                }
            }
        } else {
            for (int p = tid; p < rows0; p += NT)
                identity_row_tables(M, levels, p, bins, &s_coef, tab, cap_rows);
        }
    };

    // ---- staging (tables meanwhile), merge levels (and the store of kind-0
    // tasks)
    // S/N tasks whose last merge level fits in LDS with the circular
    // extension of its prefix sums: that level is written at the extended
    // pitch, so that every window start reads its window in place
    constexpr int kLdsFloats = kFfaShape[ffa_shape_of(NT, STAGE)].lds_bytes / 4;
    int spitch = pitch;
    int ext_w = 0;
    if (tk.kind == 1 && levels > 0) {
        const int wmax = __builtin_amdgcn_readfirstlane(tk.wmax);
        const int final_rows = tk.cone ? tk.hi - tk.lo : M;
        const int p2 = ffa_snr_pitch(bins, wmax);
        if (wmax >= 3 && wmax < bins && final_rows <= tk.ext_rows &&
            final_rows * (p2 + tk.nw) <= kLdsFloats - pitch) {
            spitch = p2;
            ext_w = wmax;
        }
    }
    rows = run_rows<NW, STAGE, V>(tk, s_lrows, data, tab, grow_v, rows, pitch, spitch, wave, lane,
                                  tables, [&]() {
                                      if (!tk.cone) setup();
                                  });
    if (rows < 0) {
        // a row length this kernel variant was not built for
        if (tid == 0) atomicOr(err, 2);
        __syncthreads();
        return;
    }
    if (tk.kind == 0) return;

    const int s_end = min(tk.hi, tk.rows_eval);
    const int nrows = max(0, s_end - tk.lo);
    // LDS past the S/N rows is free: the other staged rows, the row tables
    // and the cone segments
    float* scratch = data + (size_t)nrows * spitch;
    const int scratch_floats =
        ext_w ? kLdsFloats - pitch - nrows * spitch
              : (cap_rows - nrows) * pitch + kFfaMaxLevels * cap_rows + (2 << kMaxConeLevels) * 5;
    snr_rows<NT>(tk, data, spitch, nrows, scratch, scratch_floats, s_w, s_hb, ext_w);
}

// One workgroup per task (the task list of one pass).
template <int NT, int STAGE, int WPS, int V>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void ffa_task_kernel(
    const FfaTask* __restrict__ tasks, int ntasks, int* err)
{
    extern __shared__ float lds[];
    if ((int)blockIdx.x >= ntasks) return;
    const FfaTask tk = tasks[blockIdx.x];
    run_ffa_task<NT, STAGE, V>(tk, lds, err);
}

// Persistent batch kernel: every workgroup drains whole transforms from a
// device queue (largest first) and runs all of a transform's passes itself,
// its intermediate states in a private scratch slot, so no workgroup ever
// waits on another.  Rows written by one pass are read by the next only by
// this workgroup, ordered by the barrier that ends every task.
template <int NT, int STAGE, int WPS, int V>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPS, WPS))) void ffa_batch_kernel(
    const FfaTransform* __restrict__ trs, int ntr, int* __restrict__ queue, float* scratch,
    int64_t slot_floats, int* err)
{
    extern __shared__ float lds[];
    __shared__ int s_item;
    float* buf_a = scratch + (size_t)blockIdx.x * slot_floats;
    float* buf_b = buf_a + slot_floats / 2;
    RT_PHASE_START();
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        // uniform across the workgroup: keep the descriptor in scalar registers
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= ntr) {
            RT_PHASE(0);
            RT_PHASE_END();
            return;
        }
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
        tk.wmax = tr.wmax;
        tk.ext_rows = tr.ext_rows;
        // the transform's task list, built by the host in next_ffa_task's order
        for (int k = 0; k < tr.ntasks; ++k) {
            const FfaTaskDesc d = tr.tasks[k];
            const int src = (d.flags >> 6) & 3;
            tk.node_r0 = d.node_r0;
            tk.node_m = d.node_m;
            tk.lo = d.lo_rows & 0xFFFF;
            tk.hi = tk.lo + (d.lo_rows >> 16);
            tk.levels = d.flags & 15;
            tk.cone = (d.flags >> 4) & 1;
            tk.kind = (d.flags >> 5) & 1;
            tk.src = src == 0 ? tr.src : src == 1 ? buf_a : buf_b;
            tk.dst = ((d.flags >> 8) & 1) ? buf_b : buf_a;
            tk.rec = tr.rec ? tr.rec + (int64_t)(d.flags >> 12) * rec_words(tr.cap) : nullptr;
            RT_PHASE(0);
            run_ffa_task<NT, STAGE, V>(tk, lds, err);
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers of variant V (instantiated by ffa_variant.hip)
// ---------------------------------------------------------------------------

template <int NT, int STAGE, int WPS, int V>
static hipError_t launch_tasks_shape(const FfaTask* d_tasks, int ntasks, int* d_err,
                                     hipStream_t stream, int lds)
{
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void*)ffa_task_kernel<NT, STAGE, WPS, V>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((ffa_task_kernel<NT, STAGE, WPS, V>), dim3(ntasks), dim3(NT), lds, stream,
                       d_tasks, ntasks, d_err);
    return hipGetLastError();
}

template <int V>
hipError_t launch_ffa_tasks_variant(const FfaTask* d_tasks, int ntasks, int* d_err,
                                    hipStream_t stream, int shape)
{
    const int lds = ffa_lds_bytes(shape);
    switch (shape) {
    case 0: return launch_tasks_shape<1024, 40, 4, V>(d_tasks, ntasks, d_err, stream, lds);
    case 1: return launch_tasks_shape<512, 40, 4, V>(d_tasks, ntasks, d_err, stream, lds);
    default: return launch_tasks_shape<512, 72, 2, V>(d_tasks, ntasks, d_err, stream, lds);
    }
}

template <int NT, int STAGE, int WPS, int V>
static hipError_t launch_batch_shape(const FfaTransform* d_trs, int ntr, int* d_queue,
                                     float* d_scratch, int64_t slot_floats, int slots, int* d_err,
                                     hipStream_t stream, int lds)
{
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void*)ffa_batch_kernel<NT, STAGE, WPS, V>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((ffa_batch_kernel<NT, STAGE, WPS, V>), dim3(slots), dim3(NT), lds, stream,
                       d_trs, ntr, d_queue, d_scratch, slot_floats, d_err);
    return hipGetLastError();
}

template <int V>
hipError_t launch_ffa_batch_variant(const FfaTransform* d_trs, int ntr, int* d_queue,
                                    float* d_scratch, int64_t slot_floats, int slots, int* d_err,
                                    hipStream_t stream, int shape)
{
    const int lds = ffa_lds_bytes(shape);
    switch (shape) {
    case 0:
        return launch_batch_shape<1024, 40, 4, V>(d_trs, ntr, d_queue, d_scratch, slot_floats,
                                                  slots, d_err, stream, lds);
    case 1:
        return launch_batch_shape<512, 40, 4, V>(d_trs, ntr, d_queue, d_scratch, slot_floats,
                                                 slots, d_err, stream, lds);
    default:
        return launch_batch_shape<512, 72, 2, V>(d_trs, ntr, d_queue, d_scratch, slot_floats,
                                                 slots, d_err, stream, lds);
    }
}

}  // namespace rt
