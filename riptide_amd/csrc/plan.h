// Host-side search planner for the FFA periodogram.
//
// The planner is the single source of truth for every size, offset and
// float64 grid of a periodogram search.  Its arithmetic follows the
// reference's plan (periodogram.hpp:54-109,133-194, downsample.hpp:21-38,
// snr.hpp:49-50) and is compiled with g++ and the reference's flags so that
// the float64/float32 results are bit-identical to the reference's.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace rt {

// One downsampling step of the plan (periodogram.hpp:153-176).
struct Step {
    double f;              // downsampling factor (1.0 => raw data)
    double tau;            // sampling time after downsampling
    double pmax_samples;   // period_max / tau
    int64_t n;             // number of downsampled samples
    int64_t ds_offset;     // offset of this step's samples in the packed
                           // downsampled buffer (-1 when f == 1)
    int32_t bstart, bstop; // inclusive bins range
    int32_t first_tr, ntr; // transforms of this step
};

// One (step, bins) FFA transform (periodogram.hpp:178-199).
struct Transform {
    int32_t step;
    int32_t bins;
    int32_t rows;
    int32_t rows_eval;
    float stdnoise;
    int64_t out_row;       // first output row (period index) of this transform
};

struct SeriesPlan {
    int64_t size = 0;
    double tsamp = 0, pmin = 0, pmax = 0;
    int64_t bmin = 0, bmax = 0;
    double ds_ini = 0, ds_geo = 0;
    int64_t nds = 0;
    std::vector<Step> steps;
    std::vector<Transform> tr;
    int64_t length = 0;    // number of trial periods
    int64_t ds_total = 0;  // packed downsampled samples over all f > 1 steps
    int64_t max_n = 0;     // largest n over steps
    int64_t max_fold = 0;  // largest rows*bins over transforms
};

// Argument checks of periodogram.hpp:25-40; returns "" when valid.
std::string periodogram_check(int64_t size, double tsamp, double pmin, double pmax,
                              int64_t bmin, int64_t bmax);

// Builds the plan; returns "" on success or the reference's error message.
std::string plan_series(int64_t size, double tsamp, double pmin, double pmax, int64_t bmin,
                        int64_t bmax, SeriesPlan* out);

// Trial periods and fold bins of every row, in plan order
// (periodogram.hpp:190-194).
void plan_periods(const SeriesPlan& p, double* periods, uint32_t* foldbins);

// Boxcar height h and baseline b of snr.hpp:49-50 for `bins` phase bins and
// width w; returns (h + b, b).
void boxcar_hb(int64_t bins, int64_t w, float* hpb, float* b);

int64_t downsampled_size(int64_t n, double f);
double downsampled_variance(int64_t n, double f);

}  // namespace rt
