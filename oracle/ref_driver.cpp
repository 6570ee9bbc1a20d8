// C-ABI driver around the reference's own header-only C++ kernels, so that the
// tests can check the oracle restatement against the real reference code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile, in this container only,
// from the headers where they lie (src/riptide/cpp/*.hpp of the reference);
// the output goes to oracle/_ref/ and is never part of the product.
//
// The reference headers use memcpy / std::min / std::string / std::swap
// without including their headers, so the standard headers come first.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "kernels.hpp"
#include "block.hpp"
#include "transforms.hpp"
#include "snr.hpp"
#include "downsample.hpp"
#include "periodogram.hpp"
#include "running_median.hpp"

namespace {
thread_local std::string g_err;

template <typename F>
int guarded(F&& fn)
{
    try {
        fn();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

int ref_ffa2(const float* in, size_t rows, size_t cols, float* temp, float* out)
{
    riptide::transform(in, rows, cols, temp, out);
    return 0;
}

int ref_snr2(const float* x, size_t rows, size_t cols, const size_t* widths, size_t nw,
             float stdnoise, float* out)
{
    return guarded([&] {
        riptide::check_stdnoise(stdnoise);
        riptide::check_trial_widths(widths, nw, cols);
        riptide::snr2(riptide::ConstBlock(x, rows, cols), widths, nw, stdnoise, out);
    });
}

void ref_circular_prefix_sum(const float* x, size_t size, size_t nsum, float* out)
{
    riptide::circular_prefix_sum(x, size, nsum, out);
}

int ref_downsample(const float* in, size_t size, double f, float* out)
{
    return guarded([&] { riptide::downsample(in, size, f, out); });
}

double ref_downsampled_variance(size_t n, double f) { return riptide::downsampled_variance(n, f); }

int ref_periodogram_length(size_t size, double tsamp, double pmin, double pmax, size_t bmin,
                           size_t bmax, size_t* length)
{
    return guarded([&] {
        *length = riptide::periodogram_length(size, tsamp, pmin, pmax, bmin, bmax);
    });
}

int ref_periodogram(const float* data, size_t size, double tsamp, const size_t* widths, size_t nw,
                    double pmin, double pmax, size_t bmin, size_t bmax, double* periods,
                    uint32_t* foldbins, float* snr)
{
    return guarded([&] {
        riptide::periodogram(data, size, tsamp, widths, nw, pmin, pmax, bmin, bmax, periods,
                             foldbins, snr);
    });
}

int ref_running_median(const float* data, size_t size, size_t width, float* out)
{
    return guarded([&] { riptide::running_median<float>(data, size, width, out); });
}

}  // extern "C"
