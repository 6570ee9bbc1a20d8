/* CPU oracle for the riptide FFA periodogram path (test infrastructure only). */
#ifndef RIPTIDE_ORACLE_H
#define RIPTIDE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_EINVAL 1

const char* or_last_error(void);

void or_fused_rollback_add(const float* x, const float* y, size_t size, size_t shift, float* z);
void or_rollback(const float* x, size_t size, size_t shift, float* out);
void or_circular_prefix_sum(const float* x, size_t size, size_t nsum, float* out);

int or_ffa2(const float* in, size_t rows, size_t cols, float* temp, float* out);

int or_check_stdnoise(float stdnoise);
int or_check_trial_widths(const uint64_t* widths, size_t nw, size_t bins);
int or_snr2(const float* x, size_t rows, size_t cols, const uint64_t* widths, size_t nw,
            float stdnoise, float* out);

size_t or_downsampled_size(size_t n, double f);
double or_downsampled_variance(size_t n, double f);
int or_check_downsampling_factor(size_t n, double f);
int or_downsample(const float* in, size_t size, double f, float* out);

int or_periodogram_check(size_t size, double tsamp, double pmin, double pmax,
                         size_t bmin, size_t bmax);
size_t or_ceilshift(size_t rows, size_t cols, double pmax);
int or_periodogram_length(size_t size, double tsamp, double pmin, double pmax,
                          size_t bmin, size_t bmax, size_t* length);
int or_periodogram(const float* data, size_t size, double tsamp, const uint64_t* widths,
                   size_t nw, double pmin, double pmax, size_t bmin, size_t bmax,
                   double* periods, uint32_t* foldbins, float* snr);

int or_periodogram_f64(const float* data, size_t size, double tsamp, const uint64_t* widths,
                       size_t nw, double pmin, double pmax, size_t bmin, size_t bmax,
                       double* snr);

int or_running_median(const float* data, size_t size, size_t width, float* out);

#ifdef __cplusplus
}
#endif

#endif
