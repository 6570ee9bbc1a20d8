// Work descriptors shared by the host planner and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace rt {

// Workgroup shapes of the FFA kernels: threads and dynamic LDS bytes.  Shape
// 0 runs one 16-wavefront workgroup per CU with ~159 KiB of LDS; shape 1 two
// 8-wavefront workgroups per CU with ~79 KiB each.
// `stage` is the number of merge-level results each lane keeps in registers
// (rows per wavefront times columns per lane must fit), `waves_per_simd` the
// occupancy the kernel is compiled for (4: 128 VGPRs, 2: 256 VGPRs).
struct FfaShape {
    int threads;
    int lds_bytes;
    int stage;
    int waves_per_simd;
};
constexpr int kFfaShapes = 3;
constexpr FfaShape kFfaShape[kFfaShapes] = {
    {1024, 160 * 1024 - 1024, 40, 4},  // one 16-wave workgroup per CU
    {512, 80 * 1024 - 1024, 40, 4},    // two 8-wave workgroups per CU
    {512, 160 * 1024 - 1024, 72, 2},   // one 8-wave workgroup per CU, deep staging
};
// Largest stage over the shapes.
constexpr int kStage = 72;
// The shape of a kernel built with `threads` threads and `stage` staged
// values per lane.
constexpr int ffa_shape_of(int threads, int stage)
{
    return threads == 1024 ? 0 : (stage == 40 ? 1 : 2);
}
// Widths whose window maxima a thread of the S/N phase keeps in registers
// (more widths take a slower path).
constexpr int kSnrMaxWidths = 16;

// LDS row pitch of rows of `bins` floats: odd, so that the lanes of a
// wavefront reading one column of consecutive rows hit distinct banks.
constexpr int ffa_pitch(int bins) { return bins | 1; }
// LDS row pitch of finished rows whose circular prefix sums are extended by
// the widest trial width wmax (snr.hpp:37-41: prefix over size + wmax).
constexpr int ffa_snr_pitch(int bins, int wmax) { return (bins + wmax) | 1; }

// Rows per wavefront per merge level with k columns per lane: what the
// staging holds, at most 16 (narrow rows would otherwise need more row-table
// registers than values).
constexpr int ffa_rows_per_wave(int k, int stage)
{
    return (stage / k) < 16 ? (stage / k) : 16;
}
// Merge levels whose row tables one task keeps in LDS (all at once).
constexpr int kFfaMaxLevels = 12;
// Columns per lane staged at once by the row loads.
constexpr int kFfaLoadK = 9;
// Largest rows per workgroup (the packed row tables hold 10-bit rows).
constexpr int kFfaMaxCap = 1023;
// Largest workgroup of the FFA kernels.
constexpr int kFfaThreads = 1024;
// Wavefront width on CDNA.
constexpr int kWave = 64;
// The FFA task kernels are compiled once per variant: variant v <
// kFfaWideVariant holds the merge code of k = 2v and 2v+1 columns per lane
// (k = ceil(bins/64) <= kFfaMaxExactK), variant kFfaWideVariant the classes
// of wider rows.  Each variant holds little code, so the register allocation
// of its loops is not spoiled by the code of other row lengths; the host
// launches work grouped by variant.
constexpr int kFfaMaxExactK = 17;
constexpr int kFfaWideVariant = (kFfaMaxExactK + 1) / 2;
constexpr int kFfaVariants = kFfaWideVariant + 1;
constexpr int ffa_variant(int bins)
{
    return (bins + kWave - 1) / kWave > kFfaMaxExactK ? kFfaWideVariant
                                                      : ((bins + kWave - 1) / kWave) / 2;
}
// Largest merge-level count of a cone task (2^kMaxConeLevels bottom segments).
constexpr int kMaxConeLevels = 7;
// Register-resident S/N positions per lane of the standalone S/N kernel
// (rows up to 64*kSnrRegsMaxK bins; longer rows take the chunked path).
constexpr int kSnrRegsMaxK = 16;
// The first kSnrMaxWidths default width trials, w <- max(w + 1, int(1.5 w))
// from w = 1 (generate_width_trials with wtsp = 1.5).  A search whose widths
// are a prefix of this list takes an S/N path with compile-time offsets.
constexpr int kDefaultWidths[16] = {1, 2, 3, 4, 6, 9, 13, 19, 28, 42, 63, 94, 141, 211, 316, 474};
// Largest number of trial widths whose results one wavefront keeps in
// registers before flushing (one per lane).
constexpr int kMaxWidthsPerFlush = 64;

// One FFA work item: a tile of rows [lo, hi) of one node of a transform's
// merge tree, computed `levels` merge levels up from a source state.
//
// The merge tree of an m-row transform splits every node of M rows into a
// head of M>>1 rows and a tail of M - (M>>1) rows (block.hpp:30-43); a node's
// transform occupies its own row range.  "State at depth d" is the buffer in
// which every node at depth d (and every single-row leaf above it) holds its
// transform.  Depth-0 state is the transform itself; the folded input is the
// state at the leaf depth.
struct FfaTask {
    const float* src;   // source state rows (row r at src + r*bins)
    float* dst;         // kind 0: destination state rows
    float* snr;         // kind 1: S/N of transform row s at snr + s*nw
    const int32_t* widths;  // trial widths (nw)
    const float* hb;        // per width: (h + b, b) pairs, 2*nw floats
    int32_t bins;
    int32_t node_r0;    // node at the output depth: first row
    int32_t node_m;     // node at the output depth: number of rows
    int32_t lo, hi;     // tile, node-relative rows [lo, hi)
    int32_t levels;     // merge levels computed by this task
    int32_t cone;       // 0: whole node from the folded input; 1: cone tile
    int32_t kind;       // 0: write the tile rows to dst; 1: boxcar S/N
    int32_t rows_eval;  // kind 1: only rows s < rows_eval get S/N
    int32_t nw;
    float stdnoise;
    int32_t pad;
    // kind 1: widest trial width, and the most rows of the last merge level
    // whose circular prefix sums fit in LDS with their extension of wmax
    // values (rows at pitch ffa_snr_pitch); 0: no extension
    int32_t wmax;
    int32_t ext_rows;
    // the task's tile record (FfaRecord layout), or null: the task builds its
    // cone and row tables itself
    const int32_t* rec;
};

// A tile record: what a task would otherwise compute before and while its
// rows load, laid out in kRecWords(cap) int32 words --
//   [0]                    rows of the bottom level (the staged rows)
//   [1 + t], t <= levels   rows of merge level t (cone tasks)
//   [kRecHeader + p]       node-relative source row of staged row p (cone)
//   [kRecHeader + cap + t*cap + p]   packed row-table entry of level t, row p
// Tiles of nodes of equal size at equal depth of one transform shape share a
// record, and so do the equal transforms of every series of a batch.
constexpr int kRecHeader = 16;
constexpr int64_t rec_words(int cap) { return kRecHeader + (int64_t)cap * (1 + kFfaMaxLevels); }

// Transforms of a plan shared by at least this many series of a batch get
// tile records.
constexpr int kRecMinSeries = 2;
// Device memory the tile records of one batch may take (shapes planned
// after it is reached get none).
constexpr int64_t kRecMaxBytes = (int64_t)4 << 30;

// One tile record to build: the task fields it depends on.
struct FfaRecordSpec {
    int64_t off;        // word offset of the record
    int32_t bins, cap;
    int32_t node_m, lo, hi, levels, cone;
    int32_t pad;
};

// A task of a transform in the task list the host builds for the batch
// kernel (transforms several series share use one list): what varies
// between its tasks.
//   flags: levels (bits 0-3) | cone << 4 | kind << 5 | source (bits 6-7: 0
//   the folded input, 1 buf_a, 2 buf_b) | destination << 8 (0 buf_a, 1
//   buf_b) | record index << 12
struct FfaTaskDesc {
    int32_t node_r0;
    int32_t node_m;
    int32_t lo_rows;    // lo | (hi - lo) << 16
    int32_t flags;
};

// One whole FFA transform for the persistent batch kernel: the pass split
// (depth[0] = leaf depth of pass 0, ..., depth[npass-1] = 0) is precomputed by
// the host planner; tiles and nodes are enumerated on the device.
constexpr int kMaxPasses = 6;
struct FfaTransform {
    const float* src;       // folded input (rows*bins samples)
    float* snr;             // S/N of row 0 (row s at snr + s*nw)
    const int32_t* widths;
    const float* hb;
    int32_t bins;
    int32_t rows;
    int32_t rows_eval;
    int32_t nw;
    float stdnoise;
    int32_t cap;            // rows per workgroup
    int32_t npass;          // 1: one task computes everything
    int32_t levels0;        // npass == 1: merge levels of the whole transform
    int32_t depth[kMaxPasses];
    int32_t series;         // owning series (bookkeeping)
    int32_t shape;          // workgroup shape (kFfaShape index) the tiling assumes
    int32_t wmax;           // as FfaTask::wmax / ext_rows
    int32_t ext_rows;
    int32_t k_last;         // tile rows of the last (S/N) cone pass
    // tile records of this transform's shape (null: none): the record of
    // pass p, node size class c (node rows = (rows >> depth[p]) + c) and
    // tile k (lo = k * tile rows) is rec + (rec_first[p][c] + k) * rec_words(cap)
    const int32_t* rec;
    int32_t rec_first[kMaxPasses][2];
    // the transform's tasks in the order next_ffa_task gives them
    const FfaTaskDesc* tasks;
    int32_t ntasks;
};

// One downsampling step of one series for the packed multi-step kernel.
struct DsStep {
    double f;           // downsampling factor (> 1)
    int64_t n;          // output samples
    int64_t out_offset; // offset into the packed output
};

}  // namespace rt
