// Exact integer emulation of the few float64 operations the downsampling
// index math needs (downsample.hpp:58-73): k*f and k*f + f rounded to float64
// (round-to-nearest-even), their floor, and the float32 edge weights.
//
// A float64 factor f is passed as an integer significand and exponent,
// f = mf * 2^ef with 2^52 <= mf < 2^53, so no float64 arithmetic runs on the
// device; results are bit-identical to the reference's double arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

typedef unsigned __int128 u128;

// value = q * 2^e (q == 0 means zero)
struct Dbl {
    uint64_t q;
    int e;
};

__host__ __device__ inline int bitlen128(u128 x)
{
    const uint64_t hi = (uint64_t)(x >> 64);
    const uint64_t lo = (uint64_t)x;
    if (hi) return 128 - __builtin_clzll(hi);
    if (lo) return 64 - __builtin_clzll(lo);
    return 0;
}

// Rounds the exact value P * 2^e to a float64 (53-bit significand, RNE).
__host__ __device__ inline Dbl round53(u128 P, int e)
{
    Dbl r;
    if (P == 0) {
        r.q = 0;
        r.e = 0;
        return r;
    }
    const int nb = bitlen128(P);
    if (nb <= 53) {
        r.q = (uint64_t)P;
        r.e = e;
        return r;
    }
    const int sh = nb - 53;
    u128 q = P >> sh;
    const u128 rem = P - (q << sh);
    const u128 half = (u128)1 << (sh - 1);
    int ee = e + sh;
    if (rem > half || (rem == half && (q & 1))) {
        q += 1;
        if (q >> 53) {
            q >>= 1;
            ee += 1;
        }
    }
    r.q = (uint64_t)q;
    r.e = ee;
    return r;
}

// fl(k * f)
__host__ __device__ inline Dbl dmul_int(uint64_t k, uint64_t mf, int ef)
{
    return round53((u128)k * mf, ef);
}

// fl(a + f), a >= 0, f > 0
__host__ __device__ inline Dbl dadd(Dbl a, uint64_t mf, int ef)
{
    if (a.q == 0) {
        Dbl r;
        r.q = mf;
        r.e = ef;
        return r;
    }
    const int emin = a.e < ef ? a.e : ef;
    const u128 s = ((u128)a.q << (a.e - emin)) + ((u128)mf << (ef - emin));
    return round53(s, emin);
}

// floor of a non-negative value
__host__ __device__ inline int64_t dfloor(Dbl a)
{
    if (a.q == 0) return 0;
    if (a.e >= 0) return (int64_t)(a.q << a.e);
    if (a.e <= -64) return 0;
    return (int64_t)(a.q >> (-a.e));
}

// float32(i - a) where the float64 difference i - a is exact, 0 <= i - a
__host__ __device__ inline float dsub_to_float(int64_t i, Dbl a)
{
    if (a.q == 0) return (float)i;
    if (a.e >= 0) {
        // a is an integer
        return (float)(i - (int64_t)(a.q << a.e));
    }
    const int s = -a.e;
    const u128 num = ((u128)(uint64_t)i << s) - (u128)a.q;  // in units of 2^-s
    // num < 2^(s+2) <= 2^55: exact in uint64
    const float v = (float)(uint64_t)num;  // one rounding (RNE)
    return ldexpf(v, -s);
}

// float32(a - i) where the float64 difference a - i is exact, a >= i
__host__ __device__ inline float dsub_from_float(Dbl a, int64_t i)
{
    if (a.q == 0) return (float)(-i);
    if (a.e >= 0) return (float)((int64_t)(a.q << a.e) - i);
    const int s = -a.e;
    const u128 num = (u128)a.q - ((u128)(uint64_t)i << s);
    const float v = (float)(uint64_t)num;
    return ldexpf(v, -s);
}

// Per-output index math of downsample.hpp:58-73.
struct DsIndex {
    int64_t imin, imax;
    float wmin, wmax;
};

__host__ __device__ inline DsIndex ds_index(int64_t k, uint64_t mf, int ef, int64_t N)
{
    const Dbl start = dmul_int((uint64_t)k, mf, ef);
    const Dbl end = dadd(start, mf, ef);
    DsIndex r;
    r.imin = dfloor(start);
    const int64_t fe = dfloor(end);
    r.imax = fe < N - 1 ? fe : N - 1;
    r.wmin = dsub_to_float(r.imin + 1, start);
    r.wmax = dsub_from_float(end, r.imax);
    return r;
}

// The same index math with 64-bit integer operations, for the factors of a
// search (1 <= f < 2^10, so -52 <= ef <= -43) and k < 2^40: every value is a
// fixed-point number with F = -ef fraction bits.  fl(k*f) rounds the 128-bit
// product k*mf at bit sh = bitlen - 53 (<= 40 < F), which only moves bits
// below the binary point and carries into the integer part when the fraction
// overflows; fl(fl(k*f) + f) likewise.  Falls back to ds_index outside that
// range.
__host__ __device__ inline int bitlen64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

__host__ __device__ inline uint64_t mul64hi(uint64_t a, uint64_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((u128)a * b) >> 64);
#endif
}

// Rounds the fixed-point value (hi:lo) to 53 significant bits (RNE) in place.
__host__ __device__ inline void round53_fixed(uint64_t& hi, uint64_t& lo)
{
    const int nb = hi ? 64 + bitlen64(hi) : bitlen64(lo);
    if (nb <= 53) return;
    const int sh = nb - 53;  // < 64 for the ranges above
    const uint64_t mask = (((uint64_t)1) << sh) - 1;
    const uint64_t rem = lo & mask;
    const uint64_t half = ((uint64_t)1) << (sh - 1);
    const bool up = rem > half || (rem == half && ((lo >> sh) & 1));
    lo -= rem;
    if (up) {
        const uint64_t add = ((uint64_t)1) << sh;
        lo += add;
        if (lo < add) ++hi;
    }
}

__host__ __device__ inline DsIndex ds_index_fast(int64_t k, uint64_t mf, int ef, int64_t N)
{
    const int F = -ef;
    if (F < 43 || F > 52 || k < 0 || k >= ((int64_t)1 << 40)) return ds_index(k, mf, ef, N);
    const uint64_t fmask = (((uint64_t)1) << F) - 1;
    // start = fl(k * f)
    uint64_t hi = mul64hi((uint64_t)k, mf);
    uint64_t lo = (uint64_t)k * mf;
    round53_fixed(hi, lo);
    const int64_t i0 = (int64_t)((hi << (64 - F)) | (lo >> F));
    const uint64_t fr0 = lo & fmask;
    // end = fl(start + f)
    uint64_t lo2 = lo + mf;
    uint64_t hi2 = hi + (lo2 < mf ? 1 : 0);
    round53_fixed(hi2, lo2);
    const int64_t i1 = (int64_t)((hi2 << (64 - F)) | (lo2 >> F));
    const uint64_t fr1 = lo2 & fmask;
    DsIndex r;
    r.imin = i0;
    r.imax = i1 < N - 1 ? i1 : N - 1;
    // (imin + 1) - start = (2^F - fr0) * 2^-F, exact in float64
    r.wmin = ldexpf((float)((fmask + 1) - fr0), -F);
    if (r.imax == i1) {
        r.wmax = ldexpf((float)fr1, -F);
    } else {
        // end - imax = (i1 - imax) + fr1 * 2^-F
        const uint64_t whole = (uint64_t)(i1 - r.imax);
        if (whole < (((uint64_t)1) << (63 - F)))
            r.wmax = ldexpf((float)((whole << F) | fr1), -F);
        else
            r.wmax = ds_index(k, mf, ef, N).wmax;
    }
    return r;
}

// imin(k) = floor(fl(k*f))
__host__ __device__ inline int64_t ds_imin(int64_t k, uint64_t mf, int ef)
{
    const int F = -ef;
    if (F < 43 || F > 52 || k < 0 || k >= ((int64_t)1 << 40))
        return dfloor(dmul_int((uint64_t)k, mf, ef));
    uint64_t hi = mul64hi((uint64_t)k, mf);
    uint64_t lo = (uint64_t)k * mf;
    round53_fixed(hi, lo);
    return (int64_t)((hi << (64 - F)) | (lo >> F));
}

// The same index math in native float64 arithmetic, for k and N below 2^31:
// the float64 product and sum are each rounded once, exactly as the
// reference's doubles and the integer emulation above (the two agree bit for
// bit), at a few instructions per output.  Build with -ffp-contract=off so
// that k*f + f is not fused.  f = mf * 2^ef is exact in a double.
__host__ __device__ inline double ds_factor(uint64_t mf, int ef) { return ldexp((double)mf, ef); }

__host__ __device__ inline DsIndex ds_index_f64(int k, double f, int N)
{
    const double start = (double)k * f;
    const double end = start + f;
    const int imin = (int)start;  // start >= 0: truncation is the floor
    const int ie = (int)end;
    const int imax = ie < N - 1 ? ie : N - 1;
    DsIndex r;
    r.imin = imin;
    r.imax = imax;
    r.wmin = (float)((double)(imin + 1) - start);
    r.wmax = (float)(end - (double)imax);
    return r;
}

#ifdef __HIPCC__
// Sum of in[imin..imax] with the reference's edge weights (downsample.hpp:
// 73-78); the interior is accumulated left to right and the last term fused
// as the reference build does (acc + wmax*x -> fma).
__device__ __forceinline__ float ds_accumulate(const float* in, const DsIndex& ix)
{
    float acc = ix.wmin * in[ix.imin];
    for (int64_t i = ix.imin + 1; i < ix.imax; ++i)
        acc += in[i];
    return __fmaf_rn(ix.wmax, in[ix.imax], acc);
}
#endif

}  // namespace rt
