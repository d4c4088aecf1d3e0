// FP64 exp and tanh of the symmetry-function and MLP kernels.
//
// The device library's FP64 exp and tanh carry double-double steps and
// special-case branches (about 60 and 190 instructions on gfx950), and the
// SF passes call them for every triplet: exp(-eta (r_ij^2 + r_ik^2 + r_jk^2))
// per G4 function and the tanh^3 cutoff of r_jk. The routines here are
// branch-free, with errors of the same order as the library's (at most
// 1.0 ulp for exp and 2.2 ulp for tanh01, measured against 40-digit values):
//
// - exp_core: x = n ln2 + r (Cody-Waite, |r| <= ln2/2) and
//   m = exp(r) - 1 = r + r^2 q(r), q the degree-9 Chebyshev interpolant of
//   (exp(r) - 1 - r)/r^2 on [-ln2/2, ln2/2] (relative error 1.7e-17),
//   rounded to double.
// - exp_f64: 2^n (1 + m) by ldexp (gradual underflow to 0, overflow to inf;
//   NaN stays NaN).
// - tanh01: tanh(x) for 0 <= x <= 1 (the cutoff's argument 1 - r/rc) as
//   -e / (2 + e) with e = expm1(-2x) = 2^n m + (2^n - 1), which is m itself
//   when n = 0, so that small x loses nothing to cancellation. It shares
//   exp_core's constants, which the kernels keep in registers; a polynomial
//   of its own would add 16 more, and the kernels already run at the
//   register limit of their occupancy.
// - tanh_f64: the same for any x through |x| and the sign (at most 2.4 ulp;
//   the device library's takes 186 instructions), for the FP64 MLP.
//
// Every step is an explicit fma or a single rounded operation, so the result
// is the same in every kernel that inlines it and in a host build of this
// header.
#pragma once

namespace hdnnp {

// n with x = n ln2 + r, and m = exp(r) - 1
__host__ __device__ __forceinline__ double exp_core(double x, double &m)
{
    const double n = __builtin_rint(x * 1.4426950408889634);          // x / ln2
    double r = __builtin_fma(-n, 0.69314718055994529, x);             // ln2 rounded to double
    r = __builtin_fma(-n, 2.3190468138462996e-17, r);                 // ln2 minus that
    double q = 2.510038549551032e-08;
    q = __builtin_fma(q, r, 2.7620088445409746e-07);
    q = __builtin_fma(q, r, 2.7557268459997064e-06);
    q = __builtin_fma(q, r, 2.4801521295954376e-05);
    q = __builtin_fma(q, r, 0.00019841269863053618);
    q = __builtin_fma(q, r, 0.0013888888917213717);
    q = __builtin_fma(q, r, 0.0083333333333300615);
    q = __builtin_fma(q, r, 0.041666666666624129);
    q = __builtin_fma(q, r, 0.16666666666666669);
    q = __builtin_fma(q, r, 0.50000000000000011);
    m = __builtin_fma(r * r, q, r);
    return n;
}

__host__ __device__ __forceinline__ double exp_f64(double x)
{
    // beyond +-1000 the result is 0 or inf either way; the clamp keeps the
    // reduction exact (comparisons leave a NaN as it is)
    x = x < -1000.0 ? -1000.0 : x;
    x = x > 1000.0 ? 1000.0 : x;
    double m;
    const double n = exp_core(x, m);
    return __builtin_ldexp(1.0 + m, (int)n);
}

__host__ __device__ __forceinline__ double tanh01(double x)
{
    double m;
    const double n = exp_core(-2.0 * x, m);
    // e = 2^n (1 + m) - 1 = 2^n m + (2^n - 1), the bracket exact (n >= -3)
    const double e = __builtin_ldexp(m, (int)n) + (__builtin_ldexp(1.0, (int)n) - 1.0);
    return -e / (2.0 + e);
}

// tanh for any x, as sign(x) tanh01's formula at |x| (the MLP's activation):
// for |x| > 400, expm1(-2|x|) is -1 in double and the result +-1
__host__ __device__ __forceinline__ double tanh_f64(double x)
{
    const double a = __builtin_fabs(x);
    double m;
    const double n = exp_core(a > 400.0 ? -800.0 : -2.0 * a, m);
    const double e = __builtin_ldexp(m, (int)n) + (__builtin_ldexp(1.0, (int)n) - 1.0);
    return __builtin_copysign(-e / (2.0 + e), x);
}

}  // namespace hdnnp
