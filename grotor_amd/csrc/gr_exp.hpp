// Binary32 exponential of the cost model, in a reference form and a fast
// form that return the same binary32 value for every binary32 argument
// (checked over all 2^32 arguments on the host by
// tests/exp_exhaustive.cpp, which compiles this header unchanged).
//
// The reference form (gr_expf) stands in for CUDA's __expf in
// src/cost/full_update.cuh:11-20: Cody-Waite reduction by ln 2, Taylor
// series to degree 10 in binary64 Horner form, scaling by 2^n, one rounding
// to binary32, results below FLT_MIN flushed to +0.  Its binary64 value
// before the rounding is within about 2^-50 of e^x.
//
// The fast form reduces by ln2/32 and looks 2^(j/32) up in a table, which
// leaves a degree-5 series (relative error below 2^-44), and derives e^(x0
// + delta) from e^x0 by one multiplication for the four argument steps the
// refresh produces (1/4 and 1/2 for cells with capacity, 3/4 and 3/2 for
// cells without).  Two binary64 approximations of e^x round to the same
// binary32 unless e^x lies between them and a rounding boundary; wherever
// the fast value lies within 2^-39 of a boundary (a tie between binary32
// neighbours, or the flush threshold 2^-126) the reference form is
// evaluated instead.  The whole library is compiled with -ffp-contract=off,
// so host and device evaluate the same binary64 operations.
#pragma once

#include <hip/hip_runtime.h>

namespace grotor {

// 2^(j/32), j = 0..31, each the nearest binary64
constexpr double kExp2Frac[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0,
};

// Binary64 value of the reference form before its rounding, x in [-88, 89].
__host__ __device__ __forceinline__ double gr_exp_horner(double x) {
    const double n = __builtin_floor(x * 0x1.71547652b82fep+0 + 0.5);
    double r = x - n * 0x1.62e42fee00000p-1;
    r = r - n * 0x1.a39ef35793c76p-33;
    double p = 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(p, (int)n);
}

// The final rounding: binary32 nearest, values below 2^-126 flushed to +0.
__host__ __device__ __forceinline__ float gr_exp_round(double v) { return v < 0x1.0p-126 ? 0.0f : (float)v; }

// Reference form, binary32 -> binary32.
__host__ __device__ __forceinline__ float gr_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 89.0f) return __builtin_huge_valf();
    if (xf < -88.0f) return 0.0f;
    return gr_exp_round(gr_exp_horner((double)xf));
}

// Table form, x in [-88, 89]: n = nearest integer to x * 32/ln2,
// r = x - n * ln2/32 (|r| < 0.0109), e^r by its Taylor series to degree 5,
// times 2^((n mod 32)/32) and 2^floor(n/32).
__host__ __device__ __forceinline__ double gr_exp_table(double x) {
    const double n = __builtin_rint(x * 0x1.71547652b82fep+5);
    const int ni = (int)n;
    const double r = x - n * 0x1.62e42fefa39efp-6;
    double p = r * (1.0 / 120.0) + (1.0 / 24.0);
    p = p * r + (1.0 / 6.0);
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(kExp2Frac[ni & 31] * p, ni >> 5);
}

// Binary64 units in the last place around a binary32 tie inside which the
// fast form defers to the reference form: 2^13 units = 2^-39 relative.
constexpr int kExpTieUnits = 1 << 13;

// True when gr_exp_round may round v differently from a value within 2^-39
// of it: v near a tie (the low 29 of its 52 fraction bits near 2^28, half a
// binary32 unit) or near the flush threshold.
__host__ __device__ __forceinline__ bool gr_exp_near_tie(double v) {
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
    const int low = (int)(bits & 0x1FFFFFFFull) - (1 << 28);
    return (low < kExpTieUnits && low > -kExpTieUnits) || (v > 0x1.ffffp-127 && v < 0x1.0001p-126);
}

// gr_expf(xf) through the table form; *v receives the binary64 value that
// was rounded (0 outside [-88, 89]), the base for gr_expf_step.  The
// argument is clamped for the table and the special results selected after
// it, so that only a tie takes a branch.
__host__ __device__ __forceinline__ float gr_expf_fast(float xf, double* v) {
    const double x = __builtin_fmin(__builtin_fmax((double)xf, -88.0), 89.0);
    double t = gr_exp_table(x);
    if (gr_exp_near_tie(t)) t = gr_exp_horner(x);
    const bool inside = xf >= -88.0f && xf <= 89.0f;
    *v = inside ? t : 0.0;
    float r = inside ? gr_exp_round(t) : 0.0f;
    r = xf > 89.0f ? __builtin_huge_valf() : r;
    return xf != xf ? xf : r;
}

// gr_expf(xf) given v0 from gr_expf_fast(x0, &v0): when xf - x0 rounds to a
// step delta in {1/4, 1/2, 3/4, 3/2}, x0 + delta rounds to xf and xf - delta
// rounds to x0 (all in binary64), v0 * e^delta unless that lies near a tie;
// otherwise gr_expf_fast(xf).
__host__ __device__ __forceinline__ float gr_expf_step(float xf, float x0, double v0, double* v) {
    const double y = (double)xf, b = (double)x0;
    const double delta = y - b;
    const double e = delta == 0.25  ? 0x1.48b5e3c3e8186p+0
                     : delta == 0.5  ? 0x1.a61298e1e069cp+0
                     : delta == 0.75 ? 0x1.0ef9db467dcf8p+1
                     : delta == 1.5  ? 0x1.1ed3fe64fc541p+2
                                     : 0.0;
    const double t = v0 * e;
    const bool step = e != 0.0 && v0 != 0.0 && xf <= 89.0f && xf >= -88.0f && b + delta == y && y - delta == b &&
                      !gr_exp_near_tie(t);
    double w = t;
    float r = gr_exp_round(t);
    if (!step) r = gr_expf_fast(xf, &w);
    *v = w;
    return r;
}

}  // namespace grotor
