// gx_math.hpp -- short FP64 primitives for the field evaluator.
//
// The library's log1p is a general-purpose routine (about 125 VALU
// instructions on gfx950); the NFW halo needs log(1 + r/r_s) 13 times per
// Dopri8 step, where it was half of the Milky Way field's instruction count.
// gx_log1p follows the classical construction (argument reduction by the
// binary exponent, then the atanh series of log; e.g. Tang, "Table-driven
// implementation of the logarithm function", ACM TOMS 16 (1990), and the
// fdlibm e_log.c derivation):
//   1 + x = 2^k m,  m in [sqrt(1/2), sqrt(2)),  f = m - 1,  s = f / (2 + f),
//   log(1 + f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R(s^2)),
//   R(z) = sum_{n>=1} 2 z^n / (2n + 1)   (|s| < 0.1716, so z < 0.0295 and ten
//   terms leave a truncation below 2^-55 relative),
//   log1p(x) = k ln2_hi + [k ln2_lo + log(1 + f) + e/u],
// where u = fl(1 + x) and e = (1 + x) - u is recovered exactly by Fast2Sum.
// ln2_hi keeps the leading 32 bits of ln 2 (so k ln2_hi is exact) and
// ln2_lo = ln 2 - ln2_hi.  Accuracy: within 2 ulp for x > -1 (checked against
// the library log1p in the GPU tests).
#pragma once

#include <hip/hip_runtime.h>

namespace gx {

// 1/d from the hardware reciprocal estimate r0 (relative error below 2^-24
// on gfx950) and one cubically convergent correction: with e = 1 - d r0,
// 1/d = r0 (1 + e + e^2 + ...), and r0 (1 + e + e^2) leaves e^3 < 2^-72.
__device__ __forceinline__ double rcp(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// 1/sqrt(x) for finite x > 0 likewise: with e = 1 - x r0^2,
// 1/sqrt(x) = r0 (1 - e)^(-1/2) = r0 (1 + e/2 + 3e^2/8 + ...).
__device__ __forceinline__ double rsqrt_fast(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * r, r, 1.0);
  return fma(r, e * fma(e, 0.375, 0.5), r);
}

// n/d: the reciprocal above, then one residual correction of the quotient.
__device__ __forceinline__ double div(double n, double d) {
  const double r = rcp(d);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}

constexpr double kMinNormal = 2.2250738585072014e-308;
constexpr double kMaxFinite = 1.7976931348623157e308;

// sqrt(x) for x >= 0 or NaN: s = x / sqrt(x) from the estimate above, then
// one residual correction s + (x - s^2) / (2 sqrt(x)) (within 0.6 ulp) for
// positive normal finite x; 0 -> 0, +inf -> +inf, NaN -> NaN, and a
// subnormal x gives 0 (below 1.5e-154, which the callers treat as 0).
__device__ __forceinline__ double sqrt_fast(double x) {
  const double r = rsqrt_fast(x);
  const double s = x * r;
  const double c = fma(0.5 * r, fma(-s, s, x), s);
  return (x >= kMinNormal && x <= kMaxFinite) ? c : (x > kMaxFinite ? x : 0.0 * x);
}

// x / 6 by the product with the rounded 1/6 and one residual correction
// (the correctly rounded quotient except in rare double-rounding cases);
// non-finite x as the division gives it.
__device__ __forceinline__ double div6(double x) {
  constexpr double kSixth = 1.0 / 6.0;
  const double q = x * kSixth;
  const double c = fma(fma(-6.0, q, x), kSixth, q);
  return (q >= -kMaxFinite && q <= kMaxFinite) ? c : q;
}

// y^N by repeated squaring.
template <int N>
__device__ __forceinline__ double ipow(double y) {
  if constexpr (N == 1) {
    return y;
  } else if constexpr (N % 2 == 0) {
    const double h = ipow<N / 2>(y);
    return h * h;
  } else {
    return ipow<N - 1>(y) * y;
  }
}

// x^(1/N) for x >= 0 or NaN (within 1.25 ulp of the exact root): with x = m 2^k,
// m in [0.5, 1) and k = N j + r, 0 <= r < N, x^(1/N) = 2^j (m 2^r)^(1/N).
// The root of w = m 2^r in [0.5, 2^(N-1)) starts from the single-precision
// hardware log2/exp2 (relative error near 1e-7); two Newton steps on
// y^N = w, y <- y + (y/N)(w / y^N - 1), each square the relative error.
template <int N>
__device__ __forceinline__ double root_n(double x) {
  int k;
  const double m = frexp(x, &k);
  int j = k / N;
  int r = k - N * j;
  if (r < 0) {
    r += N;
    j -= 1;
  }
  const double w = ldexp(m, r);
  double y = (double)__builtin_amdgcn_exp2f(
      __builtin_amdgcn_logf((float)w) * (1.0f / (float)N));
#pragma unroll
  for (int it = 0; it < 2; ++it)
    y = fma(y * (1.0 / N), fma(w, rcp(ipow<N>(y)), -1.0), y);
  return (x > 0.0 && x <= kMaxFinite) ? ldexp(y, j) : x;
}

// log1p(x) given inv_u ~ 1/(1 + x) to within a few ulp (the caller often
// has it already); used only for the rounding correction of 1 + x.
__device__ __forceinline__ double log1p_inv(double x, double inv_u);

__device__ __forceinline__ double log1p_fast(double x) {
  return log1p_inv(x, rcp(1.0 + x));
}

__device__ __forceinline__ double log1p_inv(double x, double inv_u) {
  constexpr double kLn2Hi = 0x1.62e42fef00000p-1;
  constexpr double kLn2Lo = 0x1.473de6af278edp-34;
  constexpr double kSqrtHalf = 0.70710678118654752440;
  const double u = 1.0 + x;
  // rounding error of u = fl(1 + x), exact (Fast2Sum with the larger first)
  const double e = (x > 1.0) ? (1.0 - (u - x)) : (x - (u - 1.0));
  int k;
  double m = frexp(u, &k);  // u = m 2^k, m in [0.5, 1)
  if (m < kSqrtHalf) {
    m *= 2.0;
    k -= 1;
  }
  const double f = m - 1.0;
  // s enters the result only through s (f^2/2 + R), near f^3/4 against a
  // result near f, so its 1.5-ulp quotient costs under 0.1 ulp
  const double s = f * rcp(2.0 + f);
  const double z = s * s;
  // R(z) = sum_{n=1..10} 2 z^n / (2n + 1) by Estrin's scheme: the same terms
  // as Horner's rule with a dependency depth of 5 instead of 10, which is
  // what bounds a lone wave's step latency
  const double z2 = z * z, z4 = z2 * z2;
  const double p0 = fma(2.0 / 5.0, z, 2.0 / 3.0);
  const double p1 = fma(2.0 / 9.0, z, 2.0 / 7.0);
  const double p2 = fma(2.0 / 13.0, z, 2.0 / 11.0);
  const double p3 = fma(2.0 / 17.0, z, 2.0 / 15.0);
  const double p4 = fma(2.0 / 21.0, z, 2.0 / 19.0);
  const double q0 = fma(p1, z2, p0), q1 = fma(p3, z2, p2);
  const double R = fma(fma(p4, z4, q1), z4, q0) * z;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  // e/u: the whole result when x is below an ulp of 1, so keep it accurate
  const double corr = e * inv_u;
  return dk * kLn2Hi - ((hfsq - (s * (hfsq + R) + (dk * kLn2Lo + corr))) - f);
}

}  // namespace gx
