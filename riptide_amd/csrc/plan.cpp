// Host-side search planner (compiled with g++ and the reference's flags; see
// plan.h).  All float64 expressions keep the reference's operand types and
// order so that the compiler makes the same rounding decisions.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstddef>

namespace rt {

int64_t downsampled_size(int64_t n, double f)
{
    // downsample.hpp:21-24
    return (int64_t)(size_t)std::floor((size_t)n / f);
}

double downsampled_variance(int64_t n, double f)
{
    // downsample.hpp:29-38
    const double k = std::floor(f);
    const double r = f - k;
    const double x = (size_t)downsampled_size(n, f) * r;
    if (x > 1)
        return f - 1.0 / 3.0;
    return std::pow(k - 1, 2) + 2.0 / 3.0 * std::pow(x, 2) - x + 1.0;
}

std::string periodogram_check(int64_t size, double tsamp, double pmin, double pmax,
                              int64_t bmin, int64_t bmax)
{
    (void)size;
    // periodogram.hpp:25-40 (same order, same messages)
    if (!(tsamp > 0)) return "tsamp must be > 0";
    if (!(pmin > 0)) return "period_min must be > 0";
    if (!(pmax > pmin)) return "period_max must be > period_min";
    if (!(bmin > 1)) return "bins_min must be > 1";
    if (!(bmax >= bmin)) return "bins_max must be >= bins_min";
    if (!(pmin >= tsamp * (size_t)bmin)) return "Must have: period_min >= tsamp * bins_min ";
    return "";
}

static size_t ceilshift(size_t rows, size_t cols, double pmax)
{
    // periodogram.hpp:54-57
    return std::ceil(cols * (rows - 1.0) * (1.0 - cols / pmax));
}

std::string plan_series(int64_t size_, double tsamp, double pmin, double pmax, int64_t bmin_,
                        int64_t bmax_, SeriesPlan* out)
{
    std::string err = periodogram_check(size_, tsamp, pmin, pmax, bmin_, bmax_);
    if (!err.empty())
        return err;
    const size_t size = (size_t)size_;
    const size_t bins_min = (size_t)bmin_;
    const size_t bins_max = (size_t)bmax_;

    SeriesPlan& p = *out;
    p = SeriesPlan();
    p.size = size_;
    p.tsamp = tsamp;
    p.pmin = pmin;
    p.pmax = pmax;
    p.bmin = bmin_;
    p.bmax = bmax_;

    // periodogram.hpp:133-141
    const double ds_ini = pmin / (tsamp * bins_min);
    const double ds_geo = (bins_max + 1.0) / bins_min;
    const size_t nds = std::ceil(std::log(pmax / pmin) / std::log(ds_geo));
    p.ds_ini = ds_ini;
    p.ds_geo = ds_geo;
    p.nds = (int64_t)nds;

    int64_t row = 0;
    int64_t ds_off = 0;
    for (size_t ids = 0; ids < nds; ++ids) {
        // periodogram.hpp:155-176
        const double f = ds_ini * std::pow(ds_geo, ids);
        const double tau = f * tsamp;
        const double period_max_samples = pmax / tau;
        const size_t n = std::floor(size / f);
        Step st;
        st.f = f;
        st.tau = tau;
        st.pmax_samples = period_max_samples;
        st.n = (int64_t)n;
        if (f == 1) {
            st.ds_offset = -1;
        } else {
            st.ds_offset = ds_off;
            ds_off += (int64_t)n;
        }
        const size_t bstart = bins_min;
        const size_t bstop = std::min({bins_max, n, size_t(period_max_samples)});
        st.bstart = (int32_t)bstart;
        st.bstop = (int32_t)bstop;
        st.first_tr = (int32_t)p.tr.size();
        for (size_t bins = bstart; bins <= bstop; ++bins) {
            // periodogram.hpp:178-183
            const size_t rows = n / bins;
            const float stdnoise = std::sqrt(rows * downsampled_variance(size_, f));
            const double period_ceil = std::min(period_max_samples, bins + 1.0);
            const size_t rows_eval = std::min(rows, ceilshift(rows, bins, period_ceil));
            Transform t;
            t.step = (int32_t)ids;
            t.bins = (int32_t)bins;
            t.rows = (int32_t)rows;
            t.rows_eval = (int32_t)rows_eval;
            t.stdnoise = stdnoise;
            t.out_row = row;
            row += (int64_t)rows_eval;
            p.tr.push_back(t);
            p.max_fold = std::max<int64_t>(p.max_fold, (int64_t)(rows * bins));
        }
        st.ntr = (int32_t)p.tr.size() - st.first_tr;
        p.max_n = std::max<int64_t>(p.max_n, (int64_t)n);
        p.steps.push_back(st);
    }
    p.length = row;
    p.ds_total = ds_off;
    return "";
}

void plan_periods(const SeriesPlan& p, double* periods, uint32_t* foldbins)
{
    for (const Transform& t : p.tr) {
        // periodogram.hpp:190-194
        const double tau = p.steps[t.step].tau;
        const size_t bins = (size_t)t.bins;
        const size_t rows = (size_t)t.rows;
        double* pp = periods + t.out_row;
        uint32_t* fb = foldbins + t.out_row;
        for (size_t s = 0; s < (size_t)t.rows_eval; ++s) {
            pp[s] = tau * bins * bins / (bins - s / (rows - 1.0));
            fb[s] = (uint32_t)bins;
        }
    }
}

void boxcar_hb(int64_t bins_, int64_t w_, float* hpb, float* bout)
{
    // snr.hpp:49-50
    const size_t size = (size_t)bins_;
    const size_t w = (size_t)w_;
    const float h = std::sqrt((size - w) / float(size * w));
    const float b = w / float(size - w) * h;
    *hpb = h + b;
    *bout = b;
}

}  // namespace rt
