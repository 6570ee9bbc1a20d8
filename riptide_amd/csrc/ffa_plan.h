// Splits one FFA transform into LDS-sized work items (FfaTask) grouped in
// passes: pass 0 transforms whole merge-tree nodes straight from the folded
// input; every later pass computes tiles of rows a few merge levels further
// up from the previous pass's state, the last one fused with the boxcar S/N
// (or writing the transform).
#pragma once

#include <cstdint>
#include <vector>

#include "tasks.h"

namespace rt {

// ceil(log2(m)): number of merge levels of an m-row node (m >= 1).
int ffa_height(int64_t m);

// Rows of `bins` columns that one workgroup of shape `shape` can stage: the
// LDS layout (rows + three per-row tables + cone segments) and the register
// staging (rows per wavefront x columns per lane <= kStage) must both fit.
int ffa_cap_rows(int bins, int shape);

// Bytes of dynamic LDS the FFA kernels are launched with, per shape.
int ffa_lds_bytes(int shape);

// Shape used when no tuned choice exists.
int ffa_default_shape();

struct TransformSpec {
    int64_t m = 0;          // rows
    int bins = 0;           // columns
    int cap = 0;            // rows per workgroup (ffa_cap_rows)
    int shape = 0;          // workgroup shape the tasks will be launched with
    const float* src = nullptr;   // folded input rows
    float* buf_a = nullptr; // intermediate states (m*bins floats each)
    float* buf_b = nullptr;
    // final output: kind 0 writes the transform to dst; kind 1 computes S/N
    int kind = 0;
    float* dst = nullptr;
    float* snr = nullptr;
    const int32_t* widths = nullptr;
    const float* hb = nullptr;
    int nw = 0;
    float stdnoise = 1.0f;
    int rows_eval = 0;
    int wmax = 0;           // widest trial width (kind 1; 0: no prefix extension)
};

// Most rows of a finished level whose circular prefix sums, extended by
// wmax values per row, fit the LDS of workgroup shape `shape` together with
// the per-width window maxima of those rows; 0 when the extension is not
// used (no S/N, more widths than the register path holds, or wmax < 3).
int ffa_snr_ext_rows(int bins, int wmax, int nw, int shape);

// Tile rows of the last (S/N) cone pass with L merge levels: the tile of
// tile_rows(cap, L) rows, cut to ext_rows where that costs at most 1/8 of
// the tile so that every tile's prefix sums take the extension.
int64_t last_tile_rows(int cap, int L, int ext_rows);

// Appends the transform's tasks to passes[0..], growing `passes` as needed;
// returns the number of passes used, or -1 if the transform cannot be tiled
// in LDS (too few rows fit).
int plan_transform_tasks(const TransformSpec& s, std::vector<std::vector<FfaTask>>& passes);

// Depths of the pass split of an m-row transform: {d1, ..., 0} (pass 0 builds
// whole nodes of depth d1, pass p the nodes of depth depths[p]); {0} when the
// transform fits one workgroup; empty when it cannot be tiled.  At most lmax
// merge levels per cone pass.
std::vector<int> transform_pass_depths(int64_t m, int cap, int lmax = kMaxConeLevels,
                                       int ext_rows = 0);

// Row additions of an m-row FFA (multiply by the bins for float adds).
int64_t ffa_adds(int64_t m);

// Number of passes plan_transform_tasks would use (no tasks built).
int count_transform_passes(int64_t m, int bins, int cap, int lmax = kMaxConeLevels,
                           int ext_rows = 0);

}  // namespace rt
