// gx_potential.hpp -- device-side evaluation of composite potentials.
//
// Closed-form gradients and Hessians of the galax potential models
// (src/galax/potential/_src/builtin/: kepler.py:94, hernquist.py:80-82,
// miyamotonagai.py:74-78, nfw/base.py:327-338, plummer.py:123-132,
// monari2016.py:110-138), summed over components as CompositePotential does
// (base_multi.py:39-82), with TranslatedPotential's Phi(x - Delta(t))
// (xfm/translate.py:100-110).  The reference obtains gradients by autodiff of
// Phi; the forms below are the same derivatives written out.
//
// Layout: the component table lives in device memory and every lane of a
// launch reads the same entries, so the loads are uniform (scalar cache).
//
// Rotating groups: consecutive components that rotate rigidly about z at the
// same rate w (GX_FRAME_ROTATING centres and untranslated Monari bars, whose
// pattern angle is phi_b + Omega t) are evaluated in the co-rotating frame:
// the point is rotated once by -w t, each member uses its fixed co-rotating
// position or bar angle, and the summed acceleration is rotated back.  One
// sincos per evaluation serves the whole group.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "../../include/galax_amd.h"
#include "gx_math.hpp"

namespace gx {

constexpr double kTiny = 2.2250738585072014e-308;  // float64 min normal

enum : int { kGroupStart = 1, kGroupEnd = 2, kInGroup = 4 };

// Device component with derived constants:
//   KEPLER:    p0 = GM
//   HERNQUIST: p0 = GM, p1 = r_s
//   MN:        p0 = GM, p1 = a, p2 = b^2
//   NFW:       p0 = GM, p1 = r_s, p2 = 1/r_s
//   PLUMMER:   p0 = GM, p1 = r_s^2
//   MONARI:    p0 = alpha v0^2/3 (R0/Rb)^3, p1 = Rb, p2 = 1/Rb, p3 = Rb^3,
//              p4 = 1/Rb^3, p5 = 2 phi_b, p6 = Omega,
//              p7 unused; f4 = cos(2 phi_b), f5 = sin(2 phi_b)
// Frame block f: OFFSET/ROTATING centre (f0, f1, f2); ROTATING/bar group
// rate f3.
struct DComp {
  int kind;
  int frame;
  int track;
  int flags;  // kGroupStart | kGroupEnd | kInGroup
  double p[8];
  double f[6];
};

// A track of knots (kTrackKnots) as one 128-byte record per knot interval
// k (t_k <= t < t_{k+1}):
//   rec[16k + 0, 1, 2] = t_k, t_{k+1}, 1 / (t_{k+1} - t_k)
//   rec[16k + 4 + 4m .. 7 + 4m] = power-basis coefficients c0..c3 of
//   coordinate m of the cubic Hermite interpolant in s = (t - t_k)/(t_{k+1} - t_k)
// so that a lookup touches one cache line and its loads are independent.
//
// A dense-solution track (kTrackPieces) as one 192-byte piece per accepted
// step of the integrated orbit, in increasing time:
//   rec[24k + 0, 1] = the step's start time t_s and 1/h (h signed)
//   rec[24k + 2 + 7m .. 8 + 7m] = power-basis coefficients C0..C6 of
//   coordinate m of the step's dense output in th = (t - t_s)/h
// then `ncell` lookup cells of 4 doubles over [t0, t1] (cells = rec + 24n):
//   (k, b1, b2, 0): every time that maps to the cell lies in piece
//   k + (t >= b1) + (t >= b2), b1 and b2 the next two piece starts (+inf
//   past the last piece)
// then the n + 1 piece boundaries, increasing (brk).  The pieces and cells
// are contiguous (lds_doubles of them), so that a kernel can stage them in
// LDS with one copy; ncell = 0 when no cell table resolves every time.
enum : int { kTrackKnots = 0, kTrackPieces = 1 };
constexpr int kPieceDoubles = 24;
constexpr int kCellDoubles = 4;

struct DTrack {
  const double* rec;
  long long n;      // knots (records: n - 1), or pieces
  double t0, t1;    // covered interval
  double inv_h;     // knots: (n - 1) / (t1 - t0), guesses the interval;
                    // pieces: ncell / (t1 - t0), the cell of t
  double q0[3];     // position held before t0
  double q1[3];     // and after t1
  int kind;
  int ncell;
  const double* cells;
  const double* brk;
  long long lds_doubles;  // pieces + cells, when ncell > 0
};

struct DPot {
  const DComp* comps;
  const DTrack* tracks;
  const double* aux;  // per-component tables (quadrature nodes), by offset
  int n;
  int ntracks;
  // By-value copies of what the mock-stream evaluator reads on every stage
  // (its track and its Plummer row's p0, p1), filled for that composition
  // only.  Kernel arguments are read-only for the launch, so a step loop
  // keeps them in scalar registers; reads through `comps` and `tracks`
  // (device memory that the loop's own stores could alias) are repeated on
  // every trip, and a select between a track end and a looked-up position
  // whose end value came from such a read compiled to a branch around a
  // load.
  DTrack trk;
  double kp[2];
};

// ------------------------------------------------------------- tracks
// Track records live in device global memory; loads through a pointer
// marked so compile to global (not flat) loads, which the compiler can wait
// on one at a time instead of draining every outstanding memory and LDS
// operation.
typedef const __attribute__((address_space(1))) double GlobalDouble;

// Record k as 8 pairs: t_k, t_k+1 | 1/(t_k+1 - t_k), unused | then the x, y
// and z cubics' power-basis coefficients, two pairs each.
__device__ __forceinline__ void load_track_record(const double* rec,
                                                  double2 (&v)[8]) {
  GlobalDouble* r = (GlobalDouble*)rec;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    v[j].x = r[2 * j];
    v[j].y = r[2 * j + 1];
  }
}

// Position of piece k (a dense-solution track) at t: Horner in th.
__device__ __forceinline__ void piece_pos(const double* r, double t,
                                          double& x, double& y, double& z) {
  const double th = (t - r[0]) * r[1];
  double c[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double* a = r + 2 + 7 * m;
    double v = a[6];
#pragma unroll
    for (int d = 5; d >= 0; --d) v = fma(v, th, a[d]);
    c[m] = v;
  }
  x = c[0];
  y = c[1];
  z = c[2];
}

__device__ inline void track_pos(const DTrack& T, double t, double& x,
                                 double& y, double& z) {
  if (T.kind == kTrackPieces) {
    if (t <= T.t0 || t >= T.t1) {
      const double* q = t <= T.t0 ? T.q0 : T.q1;
      x = q[0];
      y = q[1];
      z = q[2];
      return;
    }
    // the last piece k with brk[k] <= t
    long long lo = 0, hi = T.n - 1;
    while (lo < hi) {
      const long long mid = (lo + hi + 1) >> 1;
      if (T.brk[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    piece_pos(T.rec + kPieceDoubles * lo, t, x, y, z);
    return;
  }
  if (T.n <= 1 || t <= T.t0) {
    x = T.q0[0];
    y = T.q0[1];
    z = T.q0[2];
    return;
  }
  if (t >= T.t1) {
    x = T.q1[0];
    y = T.q1[1];
    z = T.q1[2];
    return;
  }
  // guess from the mean spacing, then correct to the interval holding t.
  // The guess is right unless t lies within rounding of a knot, so the
  // whole 128-byte record is loaded at once (one memory round trip, where
  // checking the bounds first and loading the coefficients after took two)
  // and then checked; only a wrong guess walks to the neighbouring records
  // and loads again.
  const long long last = T.n - 2;
  long long k = (long long)floor((t - T.t0) * T.inv_h);
  k = k < 0 ? 0 : (k > last ? last : k);
  double2 v[8];
  load_track_record(T.rec + 16 * k, v);
  if (!((k == 0 || v[0].x <= t) && (k == last || v[0].y > t))) {
    const double* r = T.rec + 16 * k;
    while (k > 0 && r[0] > t) {
      --k;
      r -= 16;
    }
    while (k < last && r[1] <= t) {
      ++k;
      r += 16;
    }
    load_track_record(r, v);
  }
  // record: t_k, t_k+1, 1/(t_k+1 - t_k), unused, then the x, y, z cubics
  const double s = (t - v[0].x) * v[1].x;
  x = fma(s, fma(s, fma(s, v[3].y, v[3].x), v[2].y), v[2].x);
  y = fma(s, fma(s, fma(s, v[5].y, v[5].x), v[4].y), v[4].x);
  z = fma(s, fma(s, fma(s, v[7].y, v[7].x), v[6].y), v[6].x);
}

// Track positions at N times, the same values as track_pos gives at each
// time, with the N records' loads issued together: one memory round trip
// for the N lookups instead of N in turn.  Every record is loaded and its
// cubic evaluated unconditionally (a time outside the track reads a clamped
// record whose value is then replaced; a one-knot track has one placeholder
// record), so that no load waits behind another lookup's branch; only a
// wrong guess (t within rounding of a knot) walks and reloads.
template <int N>
__device__ __forceinline__ void track_pos_n(const DTrack& T,
                                            const double (&t)[N],
                                            double (&d)[N][3]) {
  // knots fit in 32 bits (the host refuses larger tracks)
  const int last = (int)T.n - 2;
  const double ulast = last > 0 ? (double)last : 0.0;
  int k[N];
  double2 v[N][8];
#pragma unroll
  for (int i = 0; i < N; ++i)
    k[i] = (int)fmin(fmax((t[i] - T.t0) * T.inv_h, 0.0), ulast);
#pragma unroll
  for (int i = 0; i < N; ++i)
    load_track_record(T.rec + 16 * (unsigned)k[i], v[i]);
  // Every record is consumed here, after all loads have issued: without
  // this the compiler sank a record's loads into a branch on the end
  // tests (a select between a track end and a loaded value reads as a
  // branch-worthy load), serialising that lookup behind the others.
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      asm volatile("" : "+v"(v[i][j].x));
      asm volatile("" : "+v"(v[i][j].y));
    }
  bool fix = false;
  bool bad[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    // record: t_k, t_k+1, 1/(t_k+1 - t_k), unused, then the x, y, z cubics
    const double s = (t[i] - v[i][0].x) * v[i][1].x;
    d[i][0] = fma(s, fma(s, fma(s, v[i][3].y, v[i][3].x), v[i][2].y), v[i][2].x);
    d[i][1] = fma(s, fma(s, fma(s, v[i][5].y, v[i][5].x), v[i][4].y), v[i][4].x);
    d[i][2] = fma(s, fma(s, fma(s, v[i][7].y, v[i][7].x), v[i][6].y), v[i][6].x);
    const bool lo = T.n <= 1 || t[i] <= T.t0, hi = !lo && t[i] >= T.t1;
    d[i][0] = lo ? T.q0[0] : (hi ? T.q1[0] : d[i][0]);
    d[i][1] = lo ? T.q0[1] : (hi ? T.q1[1] : d[i][1]);
    d[i][2] = lo ? T.q0[2] : (hi ? T.q1[2] : d[i][2]);
    bad[i] = !lo && !hi &&
             !((k[i] == 0 || v[i][0].x <= t[i]) && (k[i] == last || v[i][0].y > t[i]));
    fix = fix || bad[i];
  }
  if (fix) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (bad[i]) track_pos(T, t[i], d[i][0], d[i][1], d[i][2]);
  }
}

__device__ inline void frame_delta(const DPot& P, const DComp& c, double t,
                                   double& dx, double& dy, double& dz) {
  dx = dy = dz = 0.0;
  if (c.frame == GX_FRAME_OFFSET) {
    dx = c.f[0];
    dy = c.f[1];
    dz = c.f[2];
  } else if (c.frame == GX_FRAME_ROTATING) {
    double s, co;
    sincos(c.f[3] * t, &s, &co);
    dx = c.f[0] * co - c.f[1] * s;
    dy = c.f[0] * s + c.f[1] * co;
    dz = c.f[2];
  } else if (c.frame == GX_FRAME_TRACK) {
    track_pos(P.tracks[c.track], t, dx, dy, dz);
  }
}

// ------------------------------------------------------- dual numbers
// Value and three partial derivatives; forward-mode differentiation of
// gradient code (Hessians on the cold paths).
struct Dual3 {
  double v, d[3];
  __device__ static Dual3 var(double x, int i) {
    Dual3 r{x, {0.0, 0.0, 0.0}};
    r.d[i] = 1.0;
    return r;
  }
};
__device__ inline Dual3 operator+(Dual3 a, Dual3 b) {
  return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ inline Dual3 operator-(Dual3 a, Dual3 b) {
  return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ inline Dual3 operator*(Dual3 a, Dual3 b) {
  return {a.v * b.v,
          {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ inline Dual3 operator*(Dual3 a, double c) { return {a.v * c, {a.d[0] * c, a.d[1] * c, a.d[2] * c}}; }
__device__ inline Dual3 operator*(double c, Dual3 a) { return a * c; }
__device__ inline Dual3 operator+(Dual3 a, double c) { return {a.v + c, {a.d[0], a.d[1], a.d[2]}}; }
__device__ inline Dual3 operator-(Dual3 a, double c) { return {a.v - c, {a.d[0], a.d[1], a.d[2]}}; }
__device__ inline Dual3 operator/(double c, Dual3 a) {
  const double q = c / a.v, k = -q / a.v;
  return {q, {k * a.d[0], k * a.d[1], k * a.d[2]}};
}
__device__ inline Dual3 operator/(Dual3 a, Dual3 b) { return a * (1.0 / b); }
__device__ inline Dual3 sqrt(Dual3 a) {
  const double s = ::sqrt(a.v), k = 0.5 / s;
  return {s, {k * a.d[0], k * a.d[1], k * a.d[2]}};
}
__device__ inline Dual3 operator-(Dual3 a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ inline Dual3 operator+(double c, Dual3 a) { return a + c; }
__device__ inline Dual3 operator-(double c, Dual3 a) { return (-a) + c; }
__device__ inline Dual3 log1p(Dual3 a) {
  const double k = 1.0 / (1.0 + a.v);
  return {::log1p(a.v), {k * a.d[0], k * a.d[1], k * a.d[2]}};
}
using ::log1p;
using ::sqrt;  // keep the double overloads visible next to the dual ones

// Regularized lower incomplete gamma P(a, x) (a > 0, x >= 0) given
// lga = log Gamma(a): the power series for x < a + 1, else 1 - Q from the
// continued fraction (modified Lentz); Press et al., Numerical Recipes 6.2.
__device__ inline double gammap(double a, double x, double lga) {
  if (!(x > 0.0)) return 0.0;
  const double pre = exp(a * log(x) - x - lga);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int n = 0; n < 500; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * 1e-17) break;
    }
    return sum * pre;
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 500; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - pre * h;
}

// B_z(a, 0) = int_0^z u^(a-1) / (1 - u) du (0 <= z < 1, a > 0), the
// generalized NFW's enclosed mass over m at x = r/r_s, z = x/(1 + x),
// a = 3 - gamma (nfw/generalized.py:233-330, which takes it as
// z^a/a 2F1(a, 1; a + 1; z), NIST DLMF 8.17.7).  With w = 1 - z = 1/(1 + x)
// given separately (no cancellation for large x):
//   z <= 1/2:  z^a sum_n z^n / (a + n);
//   z >  1/2:  z^a sum_k (a)_k / k! [psi(k + 1) - psi(a + k) - ln w] w^k,
// the connection formula of 2F1 in the logarithmic case c = a + b (DLMF
// 15.8.10, b = 1).  psi_a = psi(a) comes from the host.  Kept out of line
// (as is gaussian_mass): inlined, the series and library calls raised the
// generic table kernels above 256 VGPRs, halving their occupancy.
__device__ __noinline__ double gnfw_mass(double a, double psi_a, double z, double w) {
  double sum;
  if (z <= 0.5) {
    double term = 1.0;
    sum = 1.0 / a;
    for (int n = 1; n < 80; ++n) {
      term *= z;
      const double t = term / (a + n);
      sum += t;
      if (t < 1e-18 * sum) break;
    }
  } else {
    constexpr double kEulerGamma = 0.57721566490153286061;
    const double lw = log(w);
    double c = 1.0, wk = 1.0, psi_k1 = -kEulerGamma, psi_ak = psi_a;
    sum = psi_k1 - psi_ak - lw;
    for (int k = 1; k < 120; ++k) {
      c *= (a + k - 1) / k;
      wk *= w;
      psi_k1 += 1.0 / k;
      psi_ak += 1.0 / (a + k - 1);
      const double t = c * wk * (psi_k1 - psi_ak - lw);
      sum += t;
      if (fabs(t) < 1e-18 * fabs(sum)) break;
    }
  }
  return pow(z, a) * sum;
}

// The Gaussian sphere's enclosed mass over m_tot at x = r/r_s:
// erf(x/sqrt2) - sqrt(2/pi) x exp(-x^2/2) (gaussian/base.py:224-235); below
// x = 1 by its power series sqrt(2/pi) sum_{n>=1} (-1)^(n+1) x^(2n+1) /
// ((2n+1) 2^(n-1) (n-1)!), where the two terms would cancel.
__device__ __noinline__ double gaussian_mass(double x) {
  constexpr double kSqrt2OverPi = 0.79788456080286535588;
  if (x < 1.0) {
    const double x2 = x * x;
    double term = x * x2, sum = 0.0;  // x^(2n+1) / (2^(n-1) (n-1)!) at n = 1
    for (int n = 1; n <= 20; ++n) {
      sum += ((n & 1) ? term : -term) / (2 * n + 1);
      term *= x2 / (2.0 * n);
    }
    return kSqrt2OverPi * sum;
  }
  return erf(x * 0.70710678118654752440) - kSqrt2OverPi * x * exp(-0.5 * x * x);
}

// --------------------------------------------------------- spherical g
// For a spherical model Phi(r): a = -(Phi'(r)/r) x.  Returns Phi'(r)/r.
__device__ __forceinline__ double sph_g(int kind, const double* p, double r2,
                                        double r, double ir) {
  switch (kind) {
    case GX_KEPLER:
      return p[0] * ir * ir * ir;
    case GX_HERNQUIST: {
      const double u = rcp(r + p[1]);
      return p[0] * u * u * ir;
    }
    case GX_NFW: {
      const double L = log1p_fast(r * p[2]);
      const double u = rcp(r + p[1]);
      return p[0] * ir * ir * (L * ir - u);
    }
    case GX_PLUMMER: {
      const double id = rsqrt_fast(r2 + p[1]);
      return p[0] * id * id * id;
    }
    case GX_ISOCHRONE: {  // Phi'/r = G m / ((b + s)^2 s), s = sqrt(r^2 + b^2)
      const double is = rsqrt_fast(r2 + p[2]);
      const double u = rcp(p[1] + (r2 + p[2]) * is);
      return p[0] * u * u * is;
    }
    case GX_JAFFE:  // Phi'/r = G m / (r^2 (r + r_s))
      return p[0] * ir * ir * rcp(r + p[1]);
    case GX_LOGARITHMIC:  // Phi'/r = v_c^2 / (r_s^2 + r^2)
      return p[0] * rcp(p[1] + r2);
    case GX_POWERLAW_CUTOFF:  // Phi'/r = G M(<r) / r^3, M(<r) / M = P(a1, r^2/r_c^2)
      return p[0] * gammap(p[1], r2 * p[2], p[3]) * ir * ir * ir;
    case GX_BURKERT: {  // G M(<r) / r^3, M(<r) = m/C (2 log1p x + log1p x^2 - 2 atan x)
      const double X = r * p[2];
      return p[0] * (2.0 * log1p(X) + log1p(X * X) - 2.0 * atan(X)) * ir * ir * ir;
    }
    case GX_STONE:  // G M(<r) / r^3, M(<r) ~ r_h atan(r/r_h) - r_c atan(r/r_c)
      return p[0] * (p[2] * atan(r / p[2]) - p[1] * atan(r / p[1])) * ir * ir * ir;
    case GX_HARMONIC_OSCILLATOR:  // Phi'/r = omega^2
      return p[0];
    case GX_HARDCUTOFF_NFW: {  // NFW inside r_t, point mass G M(<r_t) outside
      if (r <= p[4]) {
        const double L = log1p_fast(r * p[2]);
        const double u = rcp(r + p[1]);
        return p[0] * ir * ir * (L * ir - u);
      }
      return p[3] * ir * ir * ir;
    }
    case GX_GNFW: {  // G m B_z(3 - gamma, 0) / r^3
      const double w = rcp(1.0 + r * p[2]);
      return p[0] * gnfw_mass(p[3], p[4], r * p[2] * w, w) * ir * ir * ir;
    }
    case GX_GAUSSIAN:  // G m_tot M(<r)/m_tot / r^3
      return p[0] * gaussian_mass(r * p[2]) * ir * ir * ir;
    default:
      return 0.0;
  }
}

// Triaxial logarithmic halo (LMJ09) rotated by phi about z:
// Phi = v^2/2 log(r_s^2 + x^T M x), grad Phi = v^2 M x / D,
// M = R^T diag(1/q1^2, 1/q2^2, 1/q3^2) R.
__device__ __forceinline__ void lmj09_acc(const double* p, double x, double y,
                                          double z, double& ax, double& ay,
                                          double& az) {
  const double c = p[5], s = p[6];
  const double xp = x * c + y * s, yp = y * c - x * s;
  const double mx = xp * p[2], my = yp * p[3], mz = z * p[4];  // diag(1/q^2) x'
  const double k = p[0] * rcp(p[1] + xp * mx + yp * my + z * mz);
  ax -= k * (mx * c - my * s);
  ay -= k * (mx * s + my * c);
  az -= k * mz;
}

// Gradient of the Long & Murali (1992) bar, written for any scalar type T
// (double in the step loop; a dual number for the Hessian).  With x' along
// the bar (angle alpha), B = b + sqrt(z^2 + c^2), T+- = sqrt((a +- x')^2 +
// y'^2 + B^2), N = x' - a + T-, D = x' + a + T+:
// Phi = K log(N / D), K = G m / (2a), grad Phi = K (grad N / N - grad D / D).
template <class T>
__device__ __forceinline__ void longmurali_grad(const double* p, T x, T y, T z,
                                                T& gx, T& gy, T& gz) {
  const double K = p[0], a = p[1], b = p[2], c2 = p[3], ca = p[4], sa = p[5];
  const T xp = x * ca + y * sa, yp = y * ca - x * sa;
  const T S = sqrt(z * z + c2);
  const T B = S + b;
  const T Y2 = yp * yp + B * B;
  const T ap = xp + a, am = xp - a;
  const T Tp = sqrt(ap * ap + Y2), Tm = sqrt(am * am + Y2);
  const T N = am + Tm, D = ap + Tp;
  const T iN = 1.0 / N, iD = 1.0 / D, iTp = 1.0 / Tp, iTm = 1.0 / Tm;
  const T gxp = K * ((am * iTm + 1.0) * iN - (ap * iTp + 1.0) * iD);
  const T gyp = K * yp * (iTm * iN - iTp * iD);
  gz = K * (B * z / S) * (iTm * iN - iTp * iD);
  gx = gxp * ca - gyp * sa;
  gy = gxp * sa + gyp * ca;
}

__device__ __forceinline__ void longmurali_acc(const double* p, double x,
                                               double y, double z, double& ax,
                                               double& ay, double& az) {
  double gx, gy, gz;
  longmurali_grad<double>(p, x, y, z, gx, gy, gz);
  ax -= gx;
  ay -= gy;
  az -= gz;
}


__device__ __forceinline__ void mn_acc(const double* p, double x, double y,
                                       double z, double& ax, double& ay,
                                       double& az) {
  const double zb = z * z + p[2];
  const double iS = rsqrt_fast(zb);
  const double S = zb * iS;
  const double A = p[1] + S;
  const double iD = rsqrt_fast(x * x + y * y + A * A);
  const double g = p[0] * iD * iD * iD;
  ax -= g * x;
  ay -= g * y;
  az -= g * z * A * iS;
}

// Triaxial Hernquist: Phi = -G m / (r' + r_s), r'^2 = x^2 + y^2/q1^2 +
// z^2/q2^2 (+ the smallest normal double); p = {G m, r_s, 1/q1^2, 1/q2^2}.
// Returns grad Phi (T = double, or Dual3 for the Hessian).
template <class T>
__device__ __forceinline__ void thernquist_grad(const double* p, T x, T y, T z,
                                                T& gx, T& gy, T& gz) {
  const T wy = y * p[2], wz = z * p[3];
  const T r = sqrt(x * x + y * wy + z * wz + kTiny);
  const T d = r + p[1];
  const T g = p[0] / (d * d * r);
  gx = g * x;
  gy = g * wy;
  gz = g * wz;
}

// Vogelsberger et al. (2008): Phi = -G m log(1 + r~/r_s) / r~ with
// r~ = (r_a + r) r_e / (r_a + r_e), r_e^2 = x^2 + y^2/q1^2 + z^2/q2^2;
// p = {G m, r_s, 1/q1^2, 1/q2^2, r_a}.  grad Phi = Phi'(r~) grad r~.
template <class T>
__device__ __forceinline__ void vogelsberger_grad(const double* p, T x, T y,
                                                  T z, T& gx, T& gy, T& gz) {
  const double rs = p[1], ra = p[4];
  const T wy = y * p[2], wz = z * p[3];
  const T re = sqrt(x * x + y * wy + z * wz + kTiny);
  const T r = sqrt(x * x + y * y + z * z + kTiny);
  const T ie = 1.0 / (re + ra);
  const T rt = (r + ra) * re * ie;
  const T irt = 1.0 / rt;
  const T d1 = p[0] * (log1p(rt * (1.0 / rs)) * irt * irt - irt / (rt + rs));
  const T kr = d1 * re * ie / r;                    // dPhi/dr / r
  const T ke = d1 * (r + ra) * ra * ie * ie / re;   // dPhi/dr_e / r_e
  gx = (kr + ke) * x;
  gy = kr * y + ke * wy;
  gz = kr * z + ke * wz;
}

// Lee & Suto (2003): Phi = phi0 [F1(s) + E F2(s) + B F3(s)], s = r/r_s,
// E = (e_b^2 + e_c^2)/2, B = (e_b^2 y^2 + e_c^2 z^2) / (2 r^2);
// p = {phi0, 1/r_s, e_b^2, e_c^2, E}.  With L = log(1 + s):
//   F1 = -L/s, F2 = -1/(2s) + 1/s^2 + (1/s - 1/s^3) L,
//   F3 = (s^2 - 3s - 6) / (2 s^2 (1 + s)) + 3 L / s^3,
// and grad B = (e . x - 2 B x) / r^2 with e = (0, e_b^2, e_c^2).
template <class T>
__device__ __forceinline__ void leesuto_grad(const double* p, T x, T y, T z,
                                             T& gx, T& gy, T& gz) {
  const T r2 = x * x + y * y + z * z + kTiny;
  const T r = sqrt(r2);
  const T ir2 = 1.0 / r2;
  const T s = r * p[1];
  const T is = 1.0 / s;
  const T iu = 1.0 / (s + 1.0);
  const T L = log1p(s);
  const T is2 = is * is, is3 = is2 * is, is4 = is2 * is2;
  const T F1p = L * is2 - is * iu;
  const T F2p = 0.5 * is2 - 2.0 * is3 + (3.0 * is4 - is2) * L + (is - is3) * iu;
  const T F3 = (s * s - 3.0 * s - 6.0) * 0.5 * is2 * iu + 3.0 * is3 * L;
  const T F3p = (((6.0 - s) * s + 21.0) * s + 12.0) * 0.5 * is3 * iu * iu +
                3.0 * is3 * iu - 9.0 * is4 * L;
  const T B = (p[2] * (y * y) + p[3] * (z * z)) * 0.5 * ir2;
  const T Fs = F1p + p[4] * F2p + B * F3p;
  const T kr = p[0] * (Fs * p[1] / r - 2.0 * B * F3 * ir2);
  const T ka = p[0] * F3 * ir2;
  gx = kr * x;
  gy = (kr + ka * p[2]) * y;
  gz = (kr + ka * p[3]) * z;
}

// Triaxial NFW density by Chandrasekhar's integral over ellipsoidal
// shells, on a Gauss-Legendre rule: Phi = -P sum_k w_k / (r_s + u_k),
// u_k^2 = a_k x^2 + b_k y^2 + c_k z^2, so grad Phi = P sum_k w_k
// (a_k x, b_k y, c_k z) / (u_k (r_s + u_k)^2).  p = {P = G m q1 q2, r_s,
// offset of the node table in aux, nodes}; node k holds {a_k = s_k^2,
// b_k = s_k^2 / (1 + (q1^2 - 1) s_k^2), c_k (same with q2), w_k / sqrt((1 +
// (q1^2 - 1) s_k^2)(1 + (q2^2 - 1) s_k^2))}.  The node loop reads one
// uniform table (scalar loads).
__device__ __forceinline__ void tnfw_acc(const double* p, const double* aux,
                                         double x, double y, double z,
                                         double& ax, double& ay, double& az) {
  const double* tab = aux + (long long)p[2];
  const int n = (int)p[3];
  const double rs = p[1];
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int k = 0; k < n; ++k) {
    const double a = tab[4 * k], b = tab[4 * k + 1], c = tab[4 * k + 2];
    const double S = fma(a, x2, fma(b, y2, fma(c, z2, kTiny)));
    const double iu = rsqrt_fast(S);
    const double d = fma(S, iu, rs);
    const double g = tab[4 * k + 3] * iu * rcp(d * d);
    sx = fma(g, a, sx);
    sy = fma(g, b, sy);
    sz = fma(g, c, sz);
  }
  ax -= p[0] * sx * x;
  ay -= p[0] * sy * y;
  az -= p[0] * sz * z;
}

// Non-spherical, non-bar models at (x, y, z) relative to their centre;
// returns false for the spherical kinds (handled through sph_g).
__device__ __forceinline__ bool shaped_acc(int kind, const double* p,
                                           const double* aux, double x,
                                           double y, double z, double& ax,
                                           double& ay, double& az) {
  if (kind == GX_MIYAMOTO_NAGAI) {
    mn_acc(p, x, y, z, ax, ay, az);
    return true;
  }
  if (kind == GX_LMJ09_LOG) {
    lmj09_acc(p, x, y, z, ax, ay, az);
    return true;
  }
  if (kind == GX_LONGMURALI_BAR) {
    longmurali_acc(p, x, y, z, ax, ay, az);
    return true;
  }
  if (kind == GX_KUZMIN) {  // Phi = -G m / sqrt(R^2 + (|z| + r_s)^2)
    const double A = fabs(z) + p[1];
    const double iD = rsqrt_fast(x * x + y * y + A * A);
    const double g = p[0] * iD * iD * iD;
    ax -= g * x;
    ay -= g * y;
    az -= g * copysign(A, z);
    return true;
  }
  if (kind == GX_SATOH) {  // Phi = -G m T^(-1/2), T = R^2 + z^2 + a (a + 2 sqrt(z^2 + b^2))
    const double S = sqrt(z * z + p[2]);
    const double iT = rsqrt_fast(x * x + y * y + z * z + p[1] * (p[1] + 2.0 * S));
    const double g = p[0] * iT * iT * iT;
    ax -= g * x;
    ay -= g * y;
    az -= g * z * (1.0 + p[1] / S);
    return true;
  }
  if (kind == GX_TRIAXIAL_NFW) {
    tnfw_acc(p, aux, x, y, z, ax, ay, az);
    return true;
  }
  if (kind == GX_TRIAXIAL_GAUSSIAN) {
    // grad Phi_i = P x_i sum_k W_k C_ki exp(-xi_k^2/2), xi_k^2 = sum_i C_ki x_i^2 / r_s^2
    // (gaussian/triaxial.py:132-221 differentiated under the integral)
    const double* tab = aux + (long long)p[2];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < (int)p[3]; ++k) {
      const double c0 = tab[4 * k], c1 = tab[4 * k + 1], c2 = tab[4 * k + 2];
      const double e = tab[4 * k + 3] * exp(-0.5 * p[1] * (c0 * x * x + c1 * y * y + c2 * z * z));
      sx = fma(c0, e, sx);
      sy = fma(c1, e, sy);
      sz = fma(c2, e, sz);
    }
    ax -= p[0] * sx * x;
    ay -= p[0] * sy * y;
    az -= p[0] * sz * z;
    return true;
  }
  if (kind == GX_HENON_HEILES) {  // grad (R^2/2 + c (x^2 y - y^3/3)) / t_s^2
    ax -= p[1] * fma(2.0 * p[0] * x, y, x);
    ay -= p[1] * fma(p[0], (x - y) * (x + y), y);
    return true;
  }
  if (kind >= GX_TRIAXIAL_HERNQUIST && kind <= GX_VOGELSBERGER_NFW) {
    double gx, gy, gz;
    if (kind == GX_TRIAXIAL_HERNQUIST)
      thernquist_grad<double>(p, x, y, z, gx, gy, gz);
    else if (kind == GX_LEESUTO_NFW)
      leesuto_grad<double>(p, x, y, z, gx, gy, gz);
    else
      vogelsberger_grad<double>(p, x, y, z, gx, gy, gz);
    ax -= gx;
    ay -= gy;
    az -= gz;
    return true;
  }
  return false;
}

// Monari bar in a frame where its pattern angle is fixed: c2 = cos 2beta,
// s2 = sin 2beta.  Phi = P F(r) W, W = (x^2 - y^2) c2 + 2 x y s2,
// F = U(r/Rb)/r^2 (U(s) = s^3 - 2 for s < 1, -s^-3 otherwise).
__device__ __forceinline__ void bar_acc(const double* p, double c2, double s2,
                                        double x, double y, double z,
                                        double& ax, double& ay, double& az) {
  const double R2 = x * x + y * y;
  const double r2 = R2 + z * z;
  const double ir = rsqrt_fast(r2);
  const double r = r2 * ir;
  const double ir2 = ir * ir;
  const double W = (x * x - y * y) * c2 + 2.0 * x * y * s2;
  double F, G;  // G = F'(r)/r
  if (r * p[2] < 1.0) {
    F = r * p[4] - 2.0 * ir2;
    G = p[4] * ir + 4.0 * ir2 * ir2;
  } else {
    F = -p[3] * ir2 * ir2 * ir;
    G = 5.0 * p[3] * ir2 * ir2 * ir2 * ir;
  }
  const double P = p[0];
  const double WG = W * G;
  ax -= P * (F * 2.0 * (x * c2 + y * s2) + WG * x);
  ay -= P * (F * 2.0 * (x * s2 - y * c2) + WG * y);
  az -= P * (WG * z);
}

// Acceleration -grad Phi at (x, y, z, t).  FULL = false compiles only the
// origin-centred, non-rotating path (the caller guarantees the table has no
// frames and no bars); FULL = true handles every component.
template <bool FULL>
__device__ __forceinline__ void accel(const DPot& P, double x, double y,
                                      double z, double t, double& ax,
                                      double& ay, double& az) {
  ax = ay = az = 0.0;
  // radius about the origin, shared by untranslated spherical components
  const double r2o = x * x + y * y + z * z;
  const double iro = rsqrt_fast(r2o + kTiny);
  const double ro = (r2o + kTiny) * iro;
  double gcos = 1.0, gsin = 0.0, xr = x, yr = y;
  double gax = 0.0, gay = 0.0, gaz = 0.0;
  for (int i = 0; i < P.n; ++i) {
    const DComp& c = P.comps[i];
    const int kind = c.kind;
    if (!FULL) {
      if (!shaped_acc(kind, c.p, P.aux, x, y, z, ax, ay, az)) {
        const double g = sph_g(kind, c.p, r2o, ro, iro);
        ax -= g * x;
        ay -= g * y;
        az -= g * z;
      }
      continue;
    }
    const int flags = c.flags;
    if (flags & kGroupStart) {
      sincos(c.f[3] * t, &gsin, &gcos);
      xr = x * gcos + y * gsin;
      yr = -x * gsin + y * gcos;
      gax = gay = gaz = 0.0;
    }
    if (flags & kInGroup) {
      // co-rotating frame
      if (kind == GX_MONARI_BAR) {
        bar_acc(c.p, c.f[4], c.f[5], xr, yr, z, gax, gay, gaz);
      } else {
        const double dx = xr - c.f[0], dy = yr - c.f[1], dz = z - c.f[2];
        if (!shaped_acc(kind, c.p, P.aux, dx, dy, dz, gax, gay, gaz)) {
          const double r2 = dx * dx + dy * dy + dz * dz;
          const double ir = rsqrt_fast(r2 + kTiny);
          const double r = (r2 + kTiny) * ir;
          const double g = sph_g(kind, c.p, r2, r, ir);
          gax -= g * dx;
          gay -= g * dy;
          gaz -= g * dz;
        }
      }
      if (flags & kGroupEnd) {
        ax += gax * gcos - gay * gsin;
        ay += gax * gsin + gay * gcos;
        az += gaz;
      }
      continue;
    }
    if (c.frame == GX_FRAME_NONE) {
      if (shaped_acc(kind, c.p, P.aux, x, y, z, ax, ay, az)) {
      } else if (kind == GX_MONARI_BAR) {
        double s2, c2;
        sincos(c.p[5] + 2.0 * c.p[6] * t, &s2, &c2);
        bar_acc(c.p, c2, s2, x, y, z, ax, ay, az);
      } else if (kind != GX_NULL) {
        const double g = sph_g(kind, c.p, r2o, ro, iro);
        ax -= g * x;
        ay -= g * y;
        az -= g * z;
      }
    } else {
      double dx, dy, dz;
      frame_delta(P, c, t, dx, dy, dz);
      dx = x - dx;
      dy = y - dy;
      dz = z - dz;
      if (shaped_acc(kind, c.p, P.aux, dx, dy, dz, ax, ay, az)) {
      } else if (kind == GX_MONARI_BAR) {
        double s2, c2;
        sincos(c.p[5] + 2.0 * c.p[6] * t, &s2, &c2);
        bar_acc(c.p, c2, s2, dx, dy, dz, ax, ay, az);
      } else if (kind != GX_NULL) {
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ir = rsqrt_fast(r2 + kTiny);
        const double r = (r2 + kTiny) * ir;
        const double g = sph_g(kind, c.p, r2, r, ir);
        ax -= g * dx;
        ay -= g * dy;
        az -= g * dz;
      }
    }
  }
}

// Straight-line evaluator for a fixed sequence of origin-centred models:
// the kinds are template arguments, the parameters still come from the
// table.  With no loop or kind dispatch, the compiler interleaves the
// components' independent dependency chains.
template <int K>
__device__ __forceinline__ void comp_acc(const DComp& c, double x, double y,
                                         double z, double r2, double r,
                                         double ir, double& ax, double& ay,
                                         double& az) {
  if constexpr (K == GX_MIYAMOTO_NAGAI) {
    mn_acc(c.p, x, y, z, ax, ay, az);
  } else {
    const double g = sph_g(K, c.p, r2, r, ir);
    ax -= g * x;
    ay -= g * y;
    az -= g * z;
  }
}

template <int... K>
struct EvalFixed {
  static constexpr bool kLarge = false;
  static constexpr bool kFrames = false;  // origin-centred table only
  static constexpr bool kInlineFirstStep = false;
  static constexpr int kWavesPerSimd = 0;  // occupancy target (0: none)
  template <size_t... I>
  __device__ __forceinline__ static void run(const DPot& P, double x, double y,
                                             double z, double& ax, double& ay,
                                             double& az,
                                             std::index_sequence<I...>) {
    ax = ay = az = 0.0;
    const double r2 = x * x + y * y + z * z;
    const double ir = rsqrt_fast(r2 + kTiny);
    const double r = (r2 + kTiny) * ir;
    (comp_acc<K>(P.comps[I], x, y, z, r2, r, ir, ax, ay, az), ...);
  }
  __device__ __forceinline__ static void acc(const DPot& P, double x, double y,
                                             double z, double /*t*/,
                                             double& ax, double& ay,
                                             double& az) {
    run(P, x, y, z, ax, ay, az, std::make_index_sequence<sizeof...(K)>{});
  }
};

// Table-driven evaluators (any composition).
template <bool FULL>
struct EvalTable {
  static constexpr bool kLarge = FULL;
  static constexpr bool kFrames = FULL;
  static constexpr bool kInlineFirstStep = false;
  // occupancy target (0: none): the frame-less table kernels keep two waves
  // per SIMD (without the target the compiler takes one wave and the AGPRs)
  static constexpr int kWavesPerSimd = FULL ? 0 : 2;
  __device__ __forceinline__ static void acc(const DPot& P, double x, double y,
                                             double z, double t, double& ax,
                                             double& ay, double& az) {
    accel<FULL>(P, x, y, z, t, ax, ay, az);
  }
};

// galax's MilkyWayPotential: disk, halo, bulge, nucleus (milkyway.py:173-236)
// The Milky Way model (builtin/milkyway.py:173-236; rows 0-3 = Miyamoto-
// Nagai disk, NFW halo, Hernquist bulge, Hernquist nucleus) written out for
// the step loop: the three spherical components share one reciprocal,
// 1/(d_h d_b d_n), from which each 1/d_k follows by two products, and the
// halo's log1p takes 1/(1 + r/r_s) = r_s/d_h from it instead of a fourth
// reciprocal.
struct EvalMW {
  static constexpr bool kLarge = false;
  static constexpr bool kFrames = false;
  // the first step is computed in the integrator (small evaluator)
  static constexpr bool kInlineFirstStep = true;
  static constexpr int kWavesPerSimd = 2;  // occupancy target (0: none)
  __device__ __forceinline__ static void acc(const DPot& P, double x, double y,
                                             double z, double /*t*/,
                                             double& ax, double& ay,
                                             double& az) {
    const double* dk = P.comps[0].p;
    const double* hk = P.comps[1].p;
    const double* bk = P.comps[2].p;
    const double* nk = P.comps[3].p;
    const double r2 = x * x + y * y + z * z;
    const double ir = rsqrt_fast(r2 + kTiny);
    const double r = (r2 + kTiny) * ir;
    // disk
    const double zb = z * z + dk[2];
    const double iS = rsqrt_fast(zb);
    const double S = zb * iS;
    const double A = dk[1] + S;
    const double iD = rsqrt_fast(x * x + y * y + A * A);
    const double gd = dk[0] * iD * iD * iD;
    // spherical components
    const double dh = r + hk[1], db = r + bk[1], dn = r + nk[1];
    const double dbn = db * dn;
    const double R = rcp(dh * dbn);
    const double dR = dh * R;
    const double uh = dbn * R, ub = dn * dR, un = db * dR;
    const double L = log1p_inv(r * hk[2], hk[1] * uh);
    const double g =
        ir * fma(nk[0] * un, un,
                 fma(bk[0] * ub, ub, (hk[0] * ir) * fma(L, ir, -uh)));
    ax = -(gd + g) * x;
    ay = -(gd + g) * y;
    az = -(gd * A * iS + g) * z;
  }
};

using EvalMilkyWay = EvalMW;

// The Milky Way model (rows 0-3) plus one group of perturbers rotating
// rigidly at a common rate w (rows 4..n-1): Monari bars at their fixed
// co-rotating pattern angle and Plummer spheres at fixed co-rotating
// positions -- a barred Milky Way with co-rotating subhaloes.  The point is
// rotated once by -w t, the group summed, the sum rotated back.
struct EvalMilkyWayRotating {
  static constexpr bool kLarge = true;  // rolled stage loop: one copy of the group loop
  static constexpr bool kFrames = true;
  static constexpr bool kInlineFirstStep = false;
  static constexpr int kWavesPerSimd = 0;  // occupancy target (0: none)
  __device__ __forceinline__ static void acc(const DPot& P, double x, double y,
                                             double z, double t, double& ax,
                                             double& ay, double& az) {
    EvalMW::acc(P, x, y, z, t, ax, ay, az);
    double s, c;
    sincos(P.comps[4].f[3] * t, &s, &c);
    const double xr = x * c + y * s, yr = y * c - x * s;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i = 4; i < P.n; ++i) {
      const DComp& m = P.comps[i];
      if (m.kind == GX_MONARI_BAR) {
        bar_acc(m.p, m.f[4], m.f[5], xr, yr, z, gx, gy, gz);
      } else {
        const double dx = xr - m.f[0], dy = yr - m.f[1], dz = z - m.f[2];
        const double id = rsqrt_fast(dx * dx + dy * dy + dz * dz + m.p[1]);
        const double g = m.p[0] * id * id * id;
        gx -= g * dx;
        gy -= g * dy;
        gz -= g * dz;
      }
    }
    ax += gx * c - gy * s;
    ay += gx * s + gy * c;
    az += gz;
  }
};

// The mock-stream field: the Milky Way model (rows 0-3) plus a Plummer
// sphere riding a track (row 4), as MockStreamGenerator builds it with a
// progenitor potential (legacy/mockstream/mockstream_generator.py:161-275).
//
// A step's stage times are known when the step starts, so the integrator
// looks up the track at all of them up front (track_n: two batches of loads
// instead of one dependent round trip per stage) and evaluates each stage
// with its position (acc_at).  The prefetched positions take the registers
// that the stage accelerations fill later, and the batched loads take more:
// the kernel runs one wave per SIMD, which the latency-bound stream (its
// longest particle's serial steps) does not miss.
struct EvalMilkyWayTracked {
  static constexpr bool kLarge = false;
  static constexpr bool kFrames = true;
  static constexpr bool kInlineFirstStep = false;
  static constexpr bool kTrackPrefetch = true;
  static constexpr int kWavesPerSimd = 1;  // occupancy target (0: none)
  // the track and the Plummer row (G m, r_s^2) are read from the by-value
  // copies in P (P.trk, P.kp), not through P.tracks and P.comps
  template <int N>
  __device__ __forceinline__ static void track_n(const DPot& P,
                                                 const double (&t)[N],
                                                 double (&d)[N][3]) {
    track_pos_n<N>(P.trk, t, d);
  }
  __device__ __forceinline__ static void acc_at(const DPot& P, double x,
                                                double y, double z, double t,
                                                double dx, double dy, double dz,
                                                double& ax, double& ay,
                                                double& az) {
    EvalMilkyWay::acc(P, x, y, z, t, ax, ay, az);
    const double X = x - dx, Y = y - dy, Z = z - dz;
    const double id = rsqrt_fast(X * X + Y * Y + Z * Z + P.kp[1]);
    const double g = P.kp[0] * id * id * id;
    ax -= g * X;
    ay -= g * Y;
    az -= g * Z;
  }
  __device__ __forceinline__ static void acc(const DPot& P, double x, double y,
                                             double z, double t, double& ax,
                                             double& ay, double& az) {
    double dx, dy, dz;
    track_pos(P.trk, t, dx, dy, dz);
    acc_at(P, x, y, z, t, dx, dy, dz, ax, ay, az);
  }
};

// Dynamic LDS of the kernels whose evaluator stages its track there.
extern __shared__ double2 gx_lds_track[];

// The mock-stream field with the progenitor's track given as the
// progenitor's own dense solution (a kTrackPieces track: one polynomial
// piece per accepted step of its integration).  The launch stages the
// pieces and their lookup cells in LDS (stage), and every stage-time lookup
// reads LDS only: the cell of t, then the piece it names, then seven
// coefficients per coordinate.  A few hundred pieces cover a progenitor's
// orbit over Gyr, where a knot track fine enough for its cubics to match
// the dense output holds tens of thousands of records that every lookup
// fetches from device memory.  Field and end-holding as in
// EvalMilkyWayTracked.
struct EvalMilkyWayPieces : EvalMilkyWayTracked {
  static constexpr bool kLdsTrack = true;
  // every lane of the workgroup copies; the caller synchronises
  __device__ __forceinline__ static void stage(const DPot& P) {
    const double2* src = reinterpret_cast<const double2*>(P.trk.rec);
    const int m = (int)(P.trk.lds_doubles / 2);
    for (int i = threadIdx.x; i < m; i += blockDim.x) gx_lds_track[i] = src[i];
  }
  template <int N>
  __device__ __forceinline__ static void track_n(const DPot& P,
                                                 const double (&t)[N],
                                                 double (&d)[N][3]) {
    const DTrack& T = P.trk;
    const double2* pc = gx_lds_track;
    const double2* cl = gx_lds_track + (kPieceDoubles / 2) * (int)T.n;
    const double cmax = (double)(T.ncell - 1);
    // Times outside [t0, t1] are clamped to it: the first piece at t0 and
    // the last at t1 give q0 and q1 bit for bit (the host computed those
    // with the same Horner steps), so the held ends need no select -- a
    // select between them and a looked-up value compiled to a branch
    // around each lookup's reads, serialising the lookups.
    double tc[N];
    int k[N];
    auto piece_of = [&](double tt) {
      const int c = (int)fmin((tt - T.t0) * T.inv_h, cmax);
      const double2 e0 = cl[2 * c], e1 = cl[2 * c + 1];
      return (int)e0.x + (tt >= e0.y ? 1 : 0) + (tt >= e1.x ? 1 : 0);
    };
    double lo = fmin(fmax(t[0], T.t0), T.t1), hi = lo;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      tc[i] = fmin(fmax(t[i], T.t0), T.t1);
      lo = fmin(lo, tc[i]);
      hi = fmax(hi, tc[i]);
    }
    // A step is much shorter than the progenitor's: when the batch's
    // earliest and latest times share a piece on every active lane, so do
    // all its times, and one piece read serves them.
    const int klo = piece_of(lo);
    if (__all(klo == piece_of(hi))) {
      const double2* r = pc + (kPieceDoubles / 2) * klo;
      double c[kPieceDoubles];
#pragma unroll
      for (int j = 0; j < kPieceDoubles / 2; ++j) {
        const double2 v = r[j];
        c[2 * j] = v.x;
        c[2 * j + 1] = v.y;
      }
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double th = (tc[i] - c[0]) * c[1];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          double v = c[2 + 7 * m + 6];
#pragma unroll
          for (int e = 5; e >= 0; --e) v = fma(v, th, c[2 + 7 * m + e]);
          d[i][m] = v;
        }
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) k[i] = piece_of(tc[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double2* r = pc + (kPieceDoubles / 2) * k[i];
      double c[kPieceDoubles];
#pragma unroll
      for (int j = 0; j < kPieceDoubles / 2; ++j) {
        const double2 v = r[j];
        c[2 * j] = v.x;
        c[2 * j + 1] = v.y;
      }
      const double th = (tc[i] - c[0]) * c[1];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        double v = c[2 + 7 * m + 6];
#pragma unroll
        for (int e = 5; e >= 0; --e) v = fma(v, th, c[2 + 7 * m + e]);
        d[i][m] = v;
      }
    }
  }
};

// ------------------------------------- adjoint pieces (origin-centred)
// For the discrete adjoint: at relative position x and for a cotangent w of
// the acceleration a(x), add H w (H = Hessian of Phi, so da/dx = -H) to Hw
// and add (da/dtheta_k) . w to th[k] for the component's table parameters
// theta = (p0, p1, p2) as given in gx_component (G m, then the lengths).
// The raw table parameters sit in DComp::p[5..7] (see derive()).
template <int K>
__device__ __forceinline__ void comp_adjoint(const DComp& c, double x,
                                             double y, double z,
                                             const double (&w)[3],
                                             double (&Hw)[3], double* th) {
  const double GM = c.p[5];
  if constexpr (K == GX_MIYAMOTO_NAGAI) {
    const double a = c.p[6], b = c.p[7];
    const double S2 = z * z + b * b, S = sqrt(S2);
    const double A = a + S;
    const double D2 = x * x + y * y + A * A, D = sqrt(D2);
    const double iD3 = 1.0 / (D2 * D), iD5 = iD3 / D2;
    const double iS = 1.0 / S;
    const double Az = A * z * iS;
    const double hxx = GM * (iD3 - 3.0 * x * x * iD5);
    const double hyy = GM * (iD3 - 3.0 * y * y * iD5);
    const double hxy = -3.0 * GM * x * y * iD5;
    const double hxz = -3.0 * GM * x * Az * iD5;
    const double hyz = -3.0 * GM * y * Az * iD5;
    const double hzz = GM * (z * z * iD3 * iS * iS + A * b * b * iD3 * iS * iS * iS -
                             3.0 * Az * Az * iD5);
    Hw[0] += hxx * w[0] + hxy * w[1] + hxz * w[2];
    Hw[1] += hxy * w[0] + hyy * w[1] + hyz * w[2];
    Hw[2] += hxz * w[0] + hyz * w[1] + hzz * w[2];
    // da/dGM = a/GM
    th[0] -= (x * w[0] + y * w[1] + Az * w[2]) * iD3;
    // da/da_param
    th[1] -= GM * (-3.0 * A * iD5 * (x * w[0] + y * w[1]) +
                   (z * iS * iD3 - 3.0 * z * A * A * iS * iD5) * w[2]);
    // da/db
    th[2] -= GM * (-3.0 * A * b * iS * iD5 * (x * w[0] + y * w[1]) +
                   z * (b * iS * iS * iD3 - A * b * iS * iS * iS * iD3 -
                        3.0 * A * A * b * iS * iS * iD5) *
                       w[2]);
  } else {
    const double r2 = x * x + y * y + z * z + kTiny;
    const double r = sqrt(r2), ir = 1.0 / r;
    double d1, d2, drs;  // Phi', Phi'', d(Phi'/r)/d r_s
    if constexpr (K == GX_KEPLER) {
      d1 = GM * ir * ir;
      d2 = -2.0 * GM * ir * ir * ir;
      drs = 0.0;
    } else if constexpr (K == GX_HERNQUIST) {
      const double rs = c.p[6], u = 1.0 / (r + rs);
      d1 = GM * u * u;
      d2 = -2.0 * GM * u * u * u;
      drs = -2.0 * GM * u * u * u * ir;
    } else if constexpr (K == GX_NFW) {
      const double rs = c.p[6], u = 1.0 / (r + rs), L = log1p(r / rs);
      d1 = GM * (L * ir * ir - u * ir);
      d2 = GM * (u * ir * ir - 2.0 * L * ir * ir * ir + (2.0 * r + rs) * u * u * ir * ir);
      drs = -GM * u * u * ir / rs;
    } else {  // Plummer
      const double rs = c.p[6], d = r2 + rs * rs, id = 1.0 / sqrt(d);
      d1 = GM * r * id * id * id;
      d2 = GM * (rs * rs - 2.0 * r2) * id * id * id * id * id;
      drs = -3.0 * GM * rs * id * id * id * id * id;
    }
    const double g = d1 * ir;  // a = -g x
    const double xw = x * w[0] + y * w[1] + z * w[2];
    const double k = (d2 - g) * ir * ir;
    Hw[0] += g * w[0] + k * xw * x;
    Hw[1] += g * w[1] + k * xw * y;
    Hw[2] += g * w[2] + k * xw * z;
    th[0] -= (g / GM) * xw;
    if constexpr (K != GX_KEPLER) th[1] -= drs * xw;
  }
}

// ------------------------------------------------ potential and Hessian
// (not on the integration hot path: each component in the inertial frame)
__device__ inline double phi_model(const DComp& c, const double* aux,
                                   double x, double y, double z, double t) {
  const double* p = c.p;
  switch (c.kind) {
    case GX_TRIAXIAL_HERNQUIST:
      return -p[0] / (sqrt(x * x + y * y * p[2] + z * z * p[3] + kTiny) + p[1]);
    case GX_TRIAXIAL_NFW: {
      const double* tab = aux + (long long)p[2];
      double sum = 0.0;
      for (int k = 0; k < (int)p[3]; ++k) {
        const double S = tab[4 * k] * x * x + tab[4 * k + 1] * y * y + tab[4 * k + 2] * z * z;
        sum += tab[4 * k + 3] / (p[1] + sqrt(S + kTiny));
      }
      return -p[0] * sum;
    }
    case GX_LEESUTO_NFW: {
      const double r2 = x * x + y * y + z * z + kTiny, r = sqrt(r2), s = r * p[1];
      const double L = log1p(s), is = 1.0 / s;
      const double F1 = -L * is;
      const double F2 = -0.5 * is + is * is + (is - is * is * is) * L;
      const double F3 = (s * s - 3.0 * s - 6.0) / (2.0 * s * s * (1.0 + s)) + 3.0 * is * is * is * L;
      const double B = (p[2] * y * y + p[3] * z * z) * 0.5 / r2;
      return p[0] * (F1 + p[4] * F2 + B * F3);
    }
    case GX_VOGELSBERGER_NFW: {
      const double re = sqrt(x * x + y * y * p[2] + z * z * p[3] + kTiny);
      const double r = sqrt(x * x + y * y + z * z + kTiny);
      const double rt = (p[4] + r) * re / (p[4] + re);
      return -p[0] * log1p(rt / p[1]) / rt;
    }
    case GX_KEPLER:
    case GX_HERNQUIST:
    case GX_NFW:
    case GX_PLUMMER: {
      const double r2 = x * x + y * y + z * z;
      const double r = sqrt(r2 + kTiny);
      if (c.kind == GX_KEPLER) return -p[0] / r;
      if (c.kind == GX_HERNQUIST) return -p[0] / (r + p[1]);
      if (c.kind == GX_NFW) return -p[0] * log1p(r * p[2]) / r;
      return -p[0] / sqrt(r2 + kTiny + p[1]);
    }
    case GX_MIYAMOTO_NAGAI: {
      const double A = p[1] + sqrt(z * z + p[2]);
      return -p[0] / sqrt(x * x + y * y + A * A);
    }
    case GX_ISOCHRONE: {
      const double r2 = x * x + y * y + z * z + kTiny;
      return -p[0] / (p[1] + sqrt(r2 + p[2]));
    }
    case GX_JAFFE: {
      const double r = sqrt(x * x + y * y + z * z + kTiny);
      return -(p[0] / p[1]) * log1p(p[1] / r);
    }
    case GX_LOGARITHMIC:
      return 0.5 * p[0] * log(p[1] + x * x + y * y + z * z + kTiny);
    case GX_BURKERT: {
      const double X = sqrt(x * x + y * y + z * z + kTiny) * p[2], xi = 1.0 / X;
      const double br = M_PI - 2.0 * (1.0 + xi) * atan(X) + 2.0 * (1.0 + xi) * log1p(X) -
                        (1.0 - xi) * log1p(X * X);
      return -p[0] * p[2] * br;
    }
    case GX_KUZMIN: {
      const double A = fabs(z) + p[1];
      return -p[0] / sqrt(x * x + y * y + A * A);
    }
    case GX_SATOH:
      return -p[0] / sqrt(x * x + y * y + z * z + p[1] * (p[1] + 2.0 * sqrt(z * z + p[2])));
    case GX_STONE: {
      const double r2 = x * x + y * y + z * z + kTiny, r = sqrt(r2);
      const double rc = p[1], rh = p[2];
      return -p[0] * ((rh * atan(r / rh) - rc * atan(r / rc)) / r +
                      0.5 * log((r2 + rh * rh) / (r2 + rc * rc)));
    }
    case GX_HARMONIC_OSCILLATOR:
      return 0.5 * p[0] * (x * x + y * y + z * z);
    case GX_TRIAXIAL_GAUSSIAN: {
      const double* tab = aux + (long long)p[2];
      double sum = 0.0;
      for (int k = 0; k < (int)p[3]; ++k)
        sum += tab[4 * k + 3] *
               exp(-0.5 * p[1] * (tab[4 * k] * x * x + tab[4 * k + 1] * y * y + tab[4 * k + 2] * z * z));
      return -p[0] * p[4] * sum;
    }
    case GX_HENON_HEILES:
      return p[1] * (0.5 * (x * x + y * y) + p[0] * (x * x * y - y * y * y / 3.0));
    case GX_HARDCUTOFF_NFW: {  // NFW + G m/(r_s + r_t) inside r_t, -G M(<r_t)/r outside
      const double r = sqrt(x * x + y * y + z * z + kTiny);
      if (r <= p[4]) return -p[0] * log1p(r * p[2]) / r + p[0] / (p[1] + p[4]);
      return -p[3] / r;
    }
    case GX_GNFW: {  // -G m (B_z(a, 0)/r + (1 - z^(2-gamma)) / ((2 - gamma) r_s))
      const double r = sqrt(x * x + y * y + z * z + kTiny);
      const double w = 1.0 / (1.0 + r * p[2]), zz = r * p[2] * w;
      const double b = p[3] - 1.0;
      return -p[0] * (gnfw_mass(p[3], p[4], zz, w) / r + (1.0 - pow(zz, b)) * p[2] / b);
    }
    case GX_GAUSSIAN: {
      const double r = sqrt(x * x + y * y + z * z + kTiny);
      return -p[0] * erf(r * p[2] * 0.70710678118654752440) / r;
    }
    case GX_POWERLAW_CUTOFF: {
      const double r2 = x * x + y * y + z * z + kTiny, r = sqrt(r2), s2 = r2 * p[2];
      return -p[0] * (gammap(p[1], s2, p[3]) / r + p[5] * (1.0 - gammap(p[4], s2, p[6])));
    }
    case GX_LMJ09_LOG: {
      const double xp = x * p[5] + y * p[6], yp = y * p[5] - x * p[6];
      return 0.5 * p[0] * log(p[1] + xp * xp * p[2] + yp * yp * p[3] + z * z * p[4]);
    }
    case GX_LONGMURALI_BAR: {
      const double a = p[1];
      const double xp = x * p[4] + y * p[5], yp = y * p[4] - x * p[5];
      const double B = p[2] + sqrt(z * z + p[3]);
      const double Y2 = yp * yp + B * B;
      const double Tp = sqrt((a + xp) * (a + xp) + Y2), Tm = sqrt((a - xp) * (a - xp) + Y2);
      return p[0] * log((xp - a + Tm) / (xp + a + Tp));
    }
    case GX_MONARI_BAR: {
      double s2, c2;
      sincos(p[5] + 2.0 * p[6] * t, &s2, &c2);
      const double R2 = x * x + y * y, r2 = R2 + z * z;
      const double r = sqrt(r2);
      const double W = (x * x - y * y) * c2 + 2.0 * x * y * s2;
      const double F = (r * p[2] < 1.0) ? r * p[4] - 2.0 / r2
                                        : -p[3] / (r2 * r2 * r);
      return p[0] * F * W;
    }
    default:
      return 0.0;
  }
}

// Hessian of one component at (x, y, z) (relative coordinates), added to H
// (row-major 3x3).
__device__ inline void hess_model(const DComp& c, const double* aux,
                                  double x, double y, double z, double t,
                                  double* H) {
  const double* p = c.p;
  const double X[3] = {x, y, z};
  switch (c.kind) {
    case GX_TRIAXIAL_NFW: {
      // H_ij = P sum_k w_k [delta_ij c_i g(u) + c_i c_j x_i x_j g'(u)/u],
      // g = 1/(u (r_s + u)^2), g'(u)/u = -(r_s + 3u) / (u^3 (r_s + u)^3)
      const double* tab = aux + (long long)p[2];
      const double rs = p[1];
      for (int k = 0; k < (int)p[3]; ++k) {
        const double C[3] = {tab[4 * k], tab[4 * k + 1], tab[4 * k + 2]};
        const double u = sqrt(C[0] * x * x + C[1] * y * y + C[2] * z * z + kTiny);
        const double d = rs + u;
        const double w = p[0] * tab[4 * k + 3];
        const double g = 1.0 / (u * d * d);
        const double h = -(rs + 3.0 * u) / (u * u * u * d * d * d);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            H[3 * a + b] += w * ((a == b ? C[a] * g : 0.0) + C[a] * C[b] * X[a] * X[b] * h);
      }
      break;
    }
    case GX_TRIAXIAL_HERNQUIST:
    case GX_LEESUTO_NFW:
    case GX_VOGELSBERGER_NFW: {
      // forward-mode derivative of the gradient
      const Dual3 X3[3] = {Dual3::var(x, 0), Dual3::var(y, 1), Dual3::var(z, 2)};
      Dual3 g[3];
      if (c.kind == GX_TRIAXIAL_HERNQUIST)
        thernquist_grad<Dual3>(p, X3[0], X3[1], X3[2], g[0], g[1], g[2]);
      else if (c.kind == GX_LEESUTO_NFW)
        leesuto_grad<Dual3>(p, X3[0], X3[1], X3[2], g[0], g[1], g[2]);
      else
        vogelsberger_grad<Dual3>(p, X3[0], X3[1], X3[2], g[0], g[1], g[2]);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) H[3 * a + b] += g[a].d[b];
      break;
    }
    case GX_KEPLER:
    case GX_HERNQUIST:
    case GX_NFW:
    case GX_PLUMMER:
    case GX_ISOCHRONE:
    case GX_JAFFE:
    case GX_LOGARITHMIC:
    case GX_POWERLAW_CUTOFF:
    case GX_BURKERT:
    case GX_STONE:
    case GX_HARMONIC_OSCILLATOR:
    case GX_HARDCUTOFF_NFW:
    case GX_GNFW:
    case GX_GAUSSIAN: {
      // H = g I + (Phi'' - g) xx^T / r^2, g = Phi'/r
      const double r2 = x * x + y * y + z * z;
      const double rr = r2 + kTiny;
      const double r = sqrt(rr);
      double d1 = 0, d2 = 0;  // Phi', Phi''
      const double GM = p[0];
      if (c.kind == GX_KEPLER) {
        d1 = GM / rr;
        d2 = -2.0 * GM / (rr * r);
      } else if (c.kind == GX_HERNQUIST) {
        const double u = 1.0 / (r + p[1]);
        d1 = GM * u * u;
        d2 = -2.0 * GM * u * u * u;
      } else if (c.kind == GX_NFW) {
        const double rs = p[1], L = log1p(r * p[2]);
        const double u = 1.0 / (r + rs);
        d1 = GM * (L / rr - u / r);
        d2 = GM * (u / rr - 2.0 * L / (rr * r) + (2.0 * r + rs) * u * u / rr);
      } else if (c.kind == GX_ISOCHRONE) {
        const double sv = sqrt(rr + p[2]), u = 1.0 / (p[1] + sv);
        d1 = GM * r * u * u / sv;
        d2 = GM * u * u / sv * (1.0 - rr / (sv * sv) - 2.0 * rr * u / sv);
      } else if (c.kind == GX_JAFFE) {
        const double u = 1.0 / (r + p[1]);
        d1 = GM * u / r;
        d2 = -GM * (2.0 * r + p[1]) * u * u / rr;
      } else if (c.kind == GX_BURKERT) {
        const double X = r * p[2], rs = p[1];
        const double M = 2.0 * log1p(X) + log1p(X * X) - 2.0 * atan(X);  // G M(<r) / p0
        d1 = GM * M / rr;
        // M'(r) = 4 r^2 / ((r + r_s)(r^2 + r_s^2)), so G M' / r^2 = p0 4 / (...)
        d2 = GM * (4.0 / ((r + rs) * (rr + rs * rs)) - 2.0 * M / (rr * r));
      } else if (c.kind == GX_STONE) {
        const double rc = p[1], rh = p[2];
        const double M = rh * atan(r / rh) - rc * atan(r / rc);
        const double dM = rh * rh / (rh * rh + rr) - rc * rc / (rc * rc + rr);
        d1 = GM * M / rr;
        d2 = GM * (dM / rr - 2.0 * M / (rr * r));
      } else if (c.kind == GX_POWERLAW_CUTOFF) {
        const double s2 = rr * p[2], P = gammap(p[1], s2, p[3]);
        const double dP = exp((p[1] - 1.0) * log(s2) - s2 - p[3]);  // dP/d(s^2)
        d1 = GM * P / rr;
        d2 = GM * (dP * 2.0 * r * p[2] / rr - 2.0 * P / (rr * r));
      } else if (c.kind == GX_LOGARITHMIC) {
        const double D = 1.0 / (p[1] + rr);
        d1 = GM * r * D;  // GM slot holds v_c^2
        d2 = GM * (p[1] - rr) * D * D;
      } else if (c.kind == GX_HARMONIC_OSCILLATOR) {
        d1 = GM * r;  // GM slot holds omega^2
        d2 = GM;
      } else if (c.kind == GX_HARDCUTOFF_NFW && r > p[4]) {
        d1 = p[3] / rr;
        d2 = -2.0 * p[3] / (rr * r);
      } else if (c.kind == GX_HARDCUTOFF_NFW) {
        const double rs = p[1], L = log1p(r * p[2]);
        const double u = 1.0 / (r + rs);
        d1 = GM * (L / rr - u / r);
        d2 = GM * (u / rr - 2.0 * L / (rr * r) + (2.0 * r + rs) * u * u / rr);
      } else if (c.kind == GX_GNFW) {
        // M(<r) = m B(x), dB/dr = x^(a-1) (1 + x)^(-a) / r_s
        const double X = r * p[2], w = 1.0 / (1.0 + X), a = p[3];
        const double B = gnfw_mass(a, p[4], X * w, w);
        const double dB = pow(X, a - 1.0) * pow(w, a) * p[2];
        d1 = GM * B / rr;
        d2 = GM * (dB / rr - 2.0 * B / (rr * r));
      } else if (c.kind == GX_GAUSSIAN) {
        // dM/dr = m_tot sqrt(2/pi) x^2 exp(-x^2/2) / r_s
        const double X = r * p[2], M = gaussian_mass(X);
        const double dM = 0.79788456080286535588 * X * X * exp(-0.5 * X * X) * p[2];
        d1 = GM * M / rr;
        d2 = GM * (dM / rr - 2.0 * M / (rr * r));
      } else {
        const double d = rr + p[1];
        const double id = 1.0 / sqrt(d);
        d1 = GM * r * id * id * id;
        d2 = GM * (p[1] - 2.0 * rr) * id * id * id * id * id;
      }
      const double g = d1 / r;
      const double k = (d2 - g) / rr;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          H[3 * a + b] += (a == b ? g : 0.0) + k * X[a] * X[b];
      break;
    }
    case GX_MIYAMOTO_NAGAI: {
      const double GM = p[0], b2 = p[2];
      const double S2 = z * z + b2, S = sqrt(S2);
      const double A = p[1] + S;
      const double D2 = x * x + y * y + A * A;
      const double D = sqrt(D2);
      const double iD3 = 1.0 / (D2 * D), iD5 = iD3 / D2;
      const double Az = A * z / S;  // d(A^2/2)/dz
      H[0] += GM * (iD3 - 3.0 * x * x * iD5);
      H[4] += GM * (iD3 - 3.0 * y * y * iD5);
      const double hxy = -3.0 * GM * x * y * iD5;
      H[1] += hxy;
      H[3] += hxy;
      const double hxz = -3.0 * GM * x * Az * iD5;
      const double hyz = -3.0 * GM * y * Az * iD5;
      H[2] += hxz;
      H[6] += hxz;
      H[5] += hyz;
      H[7] += hyz;
      H[8] += GM * (z * z * iD3 / S2 + A * b2 * iD3 / (S2 * S) -
                    3.0 * Az * Az * iD5);
      break;
    }
    case GX_MONARI_BAR: {
      double s2, c2;
      sincos(p[5] + 2.0 * p[6] * t, &s2, &c2);
      const double R2 = x * x + y * y, r2 = R2 + z * z;
      const double r = sqrt(r2);
      const double W = (x * x - y * y) * c2 + 2.0 * x * y * s2;
      const double gW[3] = {2.0 * (x * c2 + y * s2), 2.0 * (x * s2 - y * c2),
                            0.0};
      const double HW[9] = {2.0 * c2, 2.0 * s2, 0, 2.0 * s2, -2.0 * c2,
                            0,        0,        0, 0};
      double F, F1, F2;
      if (r * p[2] < 1.0) {
        F = r * p[4] - 2.0 / r2;
        F1 = p[4] + 4.0 / (r2 * r);
        F2 = -12.0 / (r2 * r2);
      } else {
        F = -p[3] / (r2 * r2 * r);
        F1 = 5.0 * p[3] / (r2 * r2 * r2);
        F2 = -30.0 * p[3] / (r2 * r2 * r2 * r);
      }
      const double G = F1 / r;
      const double k = (F2 - G) / r2;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double HF = (a == b ? G : 0.0) + k * X[a] * X[b];
          H[3 * a + b] += p[0] * (F * HW[3 * a + b] + G * X[a] * gW[b] +
                                  gW[a] * G * X[b] + W * HF);
        }
      break;
    }
    case GX_TRIAXIAL_GAUSSIAN: {
      // H_ij = P sum_k W_k e_k (delta_ij C_ki - C_ki C_kj x_i x_j / r_s^2)
      const double* tab = aux + (long long)p[2];
      for (int k = 0; k < (int)p[3]; ++k) {
        const double C[3] = {tab[4 * k], tab[4 * k + 1], tab[4 * k + 2]};
        const double e = p[0] * tab[4 * k + 3] * exp(-0.5 * p[1] * (C[0] * x * x + C[1] * y * y + C[2] * z * z));
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            H[3 * a + b] += e * ((a == b ? C[a] : 0.0) - C[a] * C[b] * X[a] * X[b] * p[1]);
      }
      break;
    }
    case GX_HENON_HEILES: {
      // (R^2/2 + c (x^2 y - y^3/3)) / t_s^2: Phi_xx = 1 + 2 c y,
      // Phi_xy = 2 c x, Phi_yy = 1 - 2 c y (over t_s^2), z-free
      const double hxy = p[1] * 2.0 * p[0] * x;
      H[0] += p[1] * (1.0 + 2.0 * p[0] * y);
      H[1] += hxy;
      H[3] += hxy;
      H[4] += p[1] * (1.0 - 2.0 * p[0] * y);
      break;
    }
    case GX_KUZMIN:
    case GX_SATOH: {
      // Phi = -G m T^(-1/2) with T = R^2 + w(z)^2 (Kuzmin: w = |z| + r_s)
      // or T = R^2 + z^2 + a^2 + 2 a sqrt(z^2 + b^2) (Satoh); with T_z, T_zz
      // the z-derivatives: H_ij = G m (T_ij T^-3/2 / 2 - 3/4 T_i T_j T^-5/2)
      double T, Tz, Tzz;
      if (c.kind == GX_KUZMIN) {
        const double w = fabs(z) + p[1];
        T = x * x + y * y + w * w;
        Tz = 2.0 * copysign(w, z);
        Tzz = 2.0;
      } else {
        const double S = sqrt(z * z + p[2]);
        T = x * x + y * y + z * z + p[1] * (p[1] + 2.0 * S);
        Tz = 2.0 * z * (1.0 + p[1] / S);
        Tzz = 2.0 * (1.0 + p[1] / S) - 2.0 * p[1] * z * z / (S * S * S);
      }
      const double iT = 1.0 / sqrt(T), i3 = iT * iT * iT, i5 = i3 * iT * iT;
      const double Tg[3] = {2.0 * x, 2.0 * y, Tz};
      const double Th[3] = {2.0, 2.0, Tzz};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          H[3 * a + b] += p[0] * ((a == b ? 0.5 * Th[a] * i3 : 0.0) - 0.75 * Tg[a] * Tg[b] * i5);
      break;
    }
    case GX_LMJ09_LOG: {
      // H = v^2 (M / D - 2 (M x)(M x)^T / D^2)
      const double c = p[5], s = p[6], i1 = p[2], i2 = p[3], i3 = p[4];
      const double M[9] = {c * c * i1 + s * s * i2, c * s * (i1 - i2), 0.0,
                           c * s * (i1 - i2),       s * s * i1 + c * c * i2, 0.0,
                           0.0,                     0.0,                     i3};
      double Mx[3];
      for (int a = 0; a < 3; ++a) Mx[a] = M[3 * a] * x + M[3 * a + 1] * y + M[3 * a + 2] * z;
      const double D = 1.0 / (p[1] + x * Mx[0] + y * Mx[1] + z * Mx[2]);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          H[3 * a + b] += p[0] * (M[3 * a + b] * D - 2.0 * Mx[a] * Mx[b] * D * D);
      break;
    }
    case GX_LONGMURALI_BAR: {
      // forward-mode derivative of the gradient
      const Dual3 X3[3] = {Dual3::var(x, 0), Dual3::var(y, 1), Dual3::var(z, 2)};
      Dual3 g[3];
      longmurali_grad<Dual3>(p, X3[0], X3[1], X3[2], g[0], g[1], g[2]);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) H[3 * a + b] += g[a].d[b];
      break;
    }
    default:
      break;
  }
}

__device__ inline double potential(const DPot& P, double x, double y,
                                   double z, double t) {
  double s = 0.0;
  for (int i = 0; i < P.n; ++i) {
    const DComp& c = P.comps[i];
    double dx, dy, dz;
    frame_delta(P, c, t, dx, dy, dz);
    s += phi_model(c, P.aux, x - dx, y - dy, z - dz, t);
  }
  return s;
}

__device__ inline void hessian(const DPot& P, double x, double y, double z,
                               double t, double* H) {
  for (int k = 0; k < 9; ++k) H[k] = 0.0;
  for (int i = 0; i < P.n; ++i) {
    const DComp& c = P.comps[i];
    double dx, dy, dz;
    frame_delta(P, c, t, dx, dy, dz);
    hess_model(c, P.aux, x - dx, y - dy, z - dz, t, H);
  }
}

}  // namespace gx
