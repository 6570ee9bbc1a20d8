// Enumeration of the work items (FfaTask) of one whole FFA transform, the
// order in which the persistent batch kernel runs them.  Host and device
// code: the batch builds the task lists of transforms several series share
// on the host with the same function the kernel walks otherwise.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tasks.h"

namespace rt {

// Node i of depth d of an m-row merge tree (row order); false when that slot
// lies under a leaf reached above depth d.
__host__ __device__ inline bool tree_node(int m, int d, int i, int* r0, int* nm)
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

// Next work item of a transform for the persistent kernel: pass 0 = whole
// nodes of depth depth[0] into buf_a, pass p >= 1 = cone tiles of the nodes
// of depth depth[p], alternating buf_a -> buf_b -> buf_a ..., the last pass
// producing S/N.  Three-pass transforms go depth first through passes 0 and
// 1: the leaves under node j of depth depth[1], then node j's tiles, then
// node j + 1, so that the tiles read rows this workgroup has just written
// (about 1/2^depth[1] of the transform apart, instead of the whole pass)
// while they are still in the caches; then the last pass.  (p, i, lo, j) is
// the cursor (j: node of depth depth[1] of a three-pass transform); (node,
// r0, nm) caches the rows of node `node` of pass p (nm = 0: a slot under a
// leaf).
__host__ __device__ inline bool next_ffa_task(const FfaTransform& tr, const int32_t* depth,
                                              const int32_t* rec_first, int& p, int& i, int& lo,
                                              int& j, int& node, int& r0, int& nm, float* buf_a,
                                              float* buf_b, FfaTask& tk)
{
    // the tile record of (pass, node size class, tile), if the shape has records
    auto record = [&](int pass, int d, int k) -> const int32_t* {
        if (tr.rec == nullptr) return nullptr;
        const int c = nm - (tr.rows >> d);
        return tr.rec + (int64_t)(rec_first[2 * pass + c] + k) * rec_words(tr.cap);
    };
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
            tk.rec = tr.rec ? tr.rec + (int64_t)rec_first[0] * rec_words(tr.cap) : nullptr;
            p = 1;
            return true;
        }
        const int d = depth[p];
        const bool dfs = tr.npass == 3 && p < 2;
        // nodes of this pass visited now: all of them, or those under node j
        const int i_end = dfs ? (j + 1) << (d - depth[1]) : 1 << d;
        if (i >= i_end) {
            if (!dfs) {
                ++p;
                i = 0;
            } else if (p == 0) {
                p = 1;  // node j's tiles
                i = j;
            } else if (j + 1 < (1 << depth[1])) {
                ++j;  // the leaves under the next node
                p = 0;
                i = j << (depth[0] - depth[1]);
            } else {
                p = 2;  // the last pass
                i = 0;
            }
            lo = 0;
            node = -1;
            continue;
        }
        if (node != i) {
            if (!tree_node(tr.rows, d, i, &r0, &nm)) nm = 0;
            node = i;
        }
        if (nm == 0) {
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
            tk.rec = record(0, d, 0);
            ++i;
            return true;
        }
        const int L = depth[p - 1] - d;
        const bool last = (p == tr.npass - 1);
        const int K = last ? tr.k_last : tr.cap - 2 * ((1 << L) - 1);
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
        tk.hi = lo + K < nm ? lo + K : nm;
        tk.levels = L;
        tk.cone = 1;
        tk.kind = last ? 1 : 0;
        tk.rec = record(p, d, lo / K);
        lo += K;
        return true;
    }
}


}  // namespace rt
