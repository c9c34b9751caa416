// gx_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of galax's orbit hot path, used as the checker for the
// HIP library and as the CPU baseline in bench.py.  Nothing in the shipped
// package links or calls this file.
//
// What it restates (paths relative to galax's src/galax/):
//   * potentials, written as the reference writes Phi, and differentiated by
//     forward-mode automatic differentiation (a jet of value, gradient and
//     Hessian) -- the same route as AbstractPotential._gradient = jax.grad(
//     _potential) (potential/_src/base.py:170-179) and _hessian =
//     jax.hessian(_potential) (base.py:230-239):
//       Kepler          builtin/kepler.py:94
//       Hernquist       builtin/hernquist.py:80-82
//       Miyamoto-Nagai  builtin/miyamotonagai.py:74-78
//       NFW             builtin/nfw/base.py:327-338
//       Plummer         builtin/plummer.py:123-132
//       Monari bar      builtin/monari2016.py:110-138
//       and the further models named at each case of phi_model below,
//     with r = sqrt(x^2+y^2+z^2 + tiny) (potential/_src/utils.py:45-126);
//   * the composite sum of component gradients (base_multi.py:39-82);
//   * translation Phi(x - Delta(t)) (xfm/translate.py:100-110);
//   * Hamilton's equations (dynamics/_src/orbit/field_hamiltonian.py:240-249);
//   * the explicit Runge-Kutta step with embedded error, the PID step-size
//     controller, the automatic initial step, the clip-to-end rule and the
//     constant step size of diffrax 0.7.0 (third-party, not vendored; the
//     call sites are dynamics/_src/orbit/solver.py:121-141,774-812 and
//     legacy/integrator.py:161-174).  The published algorithms are restated:
//     Prince & Dormand (1981) RK8(7)13M for Dopri8, Dormand & Prince (1980)
//     RK5(4)7M for Dopri5, Hairer-Norsett-Wanner's initial-step rule and the
//     PID formula of Soderlind as parameterised by diffrax;
//   * both batch semantics: independent per-orbit solves
//     (orbit/solver.py:854-1039) and one shared solve over a (B,6) state
//     (orbit/solver.py:774-812).
// Interior save times are served by the accepted step's dense output, as
// diffrax's SaveAt(ts) does (the Dopri8 and Dopri5 interpolants are restated
// at the Tableau below).  With a PID controller that sets dtmin and no dt0,
// the first step is dtmin: the one reference fixture that sets dtmin (the
// integrate_field doctest, dynamics/_src/solver.py:342-365) is reproduced to
// its printed digits from that first step and from no automatic one.

#include "../include/galax_amd.h"

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kTiny = 2.2250738585072014e-308;  // float64 min normal

// ------------------------------------------------------------------- jets
// Value, gradient and (ORD == 2) Hessian with respect to (x, y, z).
template <int ORD>
struct J {
  double v = 0;
  double g[3] = {0, 0, 0};
  double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

template <int O>
J<O> jconst(double c) {
  J<O> r;
  r.v = c;
  return r;
}
template <int O>
J<O> jvar(double v, int i) {
  J<O> r;
  r.v = v;
  r.g[i] = 1.0;
  return r;
}
template <int O>
J<O> operator+(const J<O>& a, const J<O>& b) {
  J<O> r;
  r.v = a.v + b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i];
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j];
  return r;
}
template <int O>
J<O> operator-(const J<O>& a) {
  J<O> r;
  r.v = -a.v;
  for (int i = 0; i < 3; ++i) r.g[i] = -a.g[i];
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.h[i][j] = -a.h[i][j];
  return r;
}
template <int O>
J<O> operator-(const J<O>& a, const J<O>& b) {
  return a + (-b);
}
template <int O>
J<O> operator+(const J<O>& a, double c) {
  J<O> r = a;
  r.v = a.v + c;
  return r;
}
template <int O>
J<O> operator+(double c, const J<O>& a) {
  return a + c;
}
template <int O>
J<O> operator-(const J<O>& a, double c) {
  J<O> r = a;
  r.v = a.v - c;
  return r;
}
template <int O>
J<O> operator-(double c, const J<O>& a) {
  return (-a) + c;
}
template <int O>
J<O> operator*(const J<O>& a, double c) {
  J<O> r;
  r.v = a.v * c;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * c;
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.h[i][j] = a.h[i][j] * c;
  return r;
}
template <int O>
J<O> operator*(double c, const J<O>& a) {
  return a * c;
}
template <int O>
J<O> operator*(const J<O>& a, const J<O>& b) {
  J<O> r;
  r.v = a.v * b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        r.h[i][j] = a.v * b.h[i][j] + b.v * a.h[i][j] + a.g[i] * b.g[j] +
                    b.g[i] * a.g[j];
  return r;
}
template <int O>
J<O> operator/(const J<O>& a, const J<O>& b) {
  // q = a/b:  q' = (a' - q b')/b,  q'' = (a'' - q b'' - q'b'^T - b'q'^T)/b
  J<O> r;
  r.v = a.v / b.v;
  for (int i = 0; i < 3; ++i) r.g[i] = (a.g[i] - r.v * b.g[i]) / b.v;
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        r.h[i][j] = (a.h[i][j] - r.v * b.h[i][j] - r.g[i] * b.g[j] -
                     b.g[i] * r.g[j]) /
                    b.v;
  return r;
}
template <int O>
J<O> operator/(const J<O>& a, double c) {
  J<O> r;
  r.v = a.v / c;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] / c;
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.h[i][j] = a.h[i][j] / c;
  return r;
}
template <int O>
J<O> operator/(double c, const J<O>& b) {
  return jconst<O>(c) / b;
}
// Chain rule for a scalar function with value f, first and second derivative.
template <int O>
J<O> chain(const J<O>& a, double f, double d1, double d2) {
  J<O> r;
  r.v = f;
  for (int i = 0; i < 3; ++i) r.g[i] = d1 * a.g[i];
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        r.h[i][j] = d1 * a.h[i][j] + d2 * a.g[i] * a.g[j];
  return r;
}
template <int O>
J<O> jsqrt(const J<O>& a) {
  double s = std::sqrt(a.v);
  return chain(a, s, 0.5 / s, -0.25 / (s * a.v));
}
template <int O>
J<O> jlog1p(const J<O>& a) {
  double d = 1.0 / (1.0 + a.v);
  return chain(a, std::log1p(a.v), d, -d * d);
}
template <int O>
J<O> jlog(const J<O>& a) {
  return chain(a, std::log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v));
}

// Regularized lower incomplete gamma P(a, x) for a > 0, x >= 0: the power
// series for x < a + 1, else 1 - Q from the continued fraction (modified
// Lentz), as in Press et al., Numerical Recipes, section 6.2.
double gammap(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double lg = std::lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int n = 0; n < 1000; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (std::fabs(del) < std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(-x + a * std::log(x) - lg);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - std::exp(-x + a * std::log(x) - lg) * h;
}

// P(a, x) of a jet x: dP/dx = x^(a-1) e^(-x) / Gamma(a)
template <int O>
J<O> jgammap(double a, const J<O>& x) {
  const double lg = std::lgamma(a);
  const double d1 = std::exp((a - 1.0) * std::log(x.v) - x.v - lg);
  const double d2 = d1 * ((a - 1.0) / x.v - 1.0);
  return chain(x, gammap(a, x.v), d1, d2);
}

template <int O>
J<O> jatan(const J<O>& a) {
  const double u = 1.0 / (1.0 + a.v * a.v);
  return chain(a, std::atan(a.v), u, -2.0 * a.v * u * u);
}

template <int O>
J<O> jabs(const J<O>& a) {  // d|x|/dx = sign(x)
  const double s = (a.v > 0) - (a.v < 0);
  return chain(a, std::fabs(a.v), s, 0.0);
}

template <int O>
J<O> jcos(const J<O>& a) {
  double c = std::cos(a.v), s = std::sin(a.v);
  return chain(a, c, -s, -c);
}
// B(x) = B_z(a, 0) = int_0^z u^(a-1) / (1 - u) du with z = x / (1 + x): the
// generalized NFW's enclosed mass over m (nfw/generalized.py:233-330).  By
// double-exponential (tanh-sinh) quadrature in long double, u = z (1 +
// tanh(pi/2 sinh s)) / 2, step 1/64 over |s| <= 4 (a different route from
// the device's series, so that each checks the other).
double gnfw_mass_q(double a, double x) {
  const long double z = (long double)x / (1.0L + x);
  const long double h = 1.0L / 64, pi2 = 1.5707963267948966192L;
  long double sum = 0.0L;
  for (int k = -256; k <= 256; ++k) {
    const long double s = k * h;
    const long double sh = std::sinh(s), ch = std::cosh(s);
    const long double th = std::tanh(pi2 * sh);
    const long double cw = std::cosh(pi2 * sh);
    const long double u = 0.5L * z * (1.0L + th);
    const long double du = 0.5L * z * pi2 * ch / (cw * cw);
    // 1 - u = (1 - z) + z (1 - th)/2, with 1 - th = 2 / (1 + exp(pi sh))
    const long double omu = 1.0L / (1.0L + x) + z / (1.0L + std::exp(2.0L * pi2 * sh));
    if (u <= 0.0L) continue;
    sum += du * std::pow(u, (long double)a - 1.0L) / omu;
  }
  return (double)(h * sum);
}
// B(X) for a jet X, with B'(x) = x^(a-1) (1 + x)^(-a) and
// B''(x) = x^(a-2) (1 + x)^(-a-1) (a - 1 - x).
template <int O>
J<O> jgnfw_mass(double a, const J<O>& X) {
  const double x = X.v;
  const double d1 = std::pow(x, a - 1.0) * std::pow(1.0 + x, -a);
  const double d2 = std::pow(x, a - 2.0) * std::pow(1.0 + x, -a - 1.0) * (a - 1.0 - x);
  return chain(X, gnfw_mass_q(a, x), d1, d2);
}
template <int O>
J<O> jpow(const J<O>& a, double b) {
  return chain(a, std::pow(a.v, b), b * std::pow(a.v, b - 1.0), b * (b - 1.0) * std::pow(a.v, b - 2.0));
}
template <int O>
J<O> jexp(const J<O>& a) {
  const double e = std::exp(a.v);
  return chain(a, e, e, e);
}
template <int O>
J<O> jerf(const J<O>& a) {
  const double d1 = 1.1283791670955125739 * std::exp(-a.v * a.v);
  return chain(a, std::erf(a.v), d1, -2.0 * a.v * d1);
}
// erf(q^(1/2)) / q^(1/2) for small q as (2/sqrt(pi)) S(q), S(q) =
// sum_m t_m / (2m + 1) with t_m = (-q)^m / m!; then S'(q) = -sum_m t_m /
// (2m + 3) and S''(q) = sum_m t_m / (2m + 5) (term-by-term derivatives).
template <int O>
J<O> jerf_over_sqrt(const J<O>& q) {
  double f = 0, d1 = 0, d2 = 0, t = 1.0;
  for (int m = 0; m < 40; ++m) {
    f += t / (2 * m + 1);
    d1 -= t / (2 * m + 3);
    d2 += t / (2 * m + 5);
    t *= -q.v / (m + 1);
  }
  const double c = 1.1283791670955125739;
  return chain(q, c * f, c * d1, c * d2);
}

template <int O>
J<O> jatan2(const J<O>& y, const J<O>& x) {
  // F(y, x) = atan2(y, x); Fy = x/r2, Fx = -y/r2,
  // Fyy = -2xy/r2^2, Fxx = 2xy/r2^2, Fxy = (y^2 - x^2)/r2^2.
  double r2 = x.v * x.v + y.v * y.v;
  double Fy = x.v / r2, Fx = -y.v / r2;
  double Fyy = -2 * x.v * y.v / (r2 * r2), Fxx = -Fyy;
  double Fxy = (y.v * y.v - x.v * x.v) / (r2 * r2);
  J<O> r;
  r.v = std::atan2(y.v, x.v);
  for (int i = 0; i < 3; ++i) r.g[i] = Fy * y.g[i] + Fx * x.g[i];
  if (O == 2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        r.h[i][j] = Fy * y.h[i][j] + Fx * x.h[i][j] + Fyy * y.g[i] * y.g[j] +
                    Fxx * x.g[i] * x.g[j] +
                    Fxy * (y.g[i] * x.g[j] + x.g[i] * y.g[j]);
  return r;
}
template <int O>
J<O> sq(const J<O>& a) {  // jnp x**2 (integer_pow)
  return a * a;
}
template <int O>
J<O> cube(const J<O>& a) {  // x**3
  return a * a * a;
}

// ------------------------------------------------------- potential models
// Gauss-Legendre rule of order n mapped to [0, 1] (nodes s_k, weights w_k):
// the roots of P_n by Newton's method on the three-term recurrence, from
// the asymptotic guesses cos(pi (k - 1/4) / (n + 1/2)); weights
// 2 / ((1 - x^2) P_n'(x)^2).  The reference's GaussLegendreIntegrator
// (potential/_src/utils.py:131-162) takes numpy's leggauss and maps it the
// same way.  Cached per order.
const std::vector<double>& gauss_legendre01(int n) {
  static std::vector<double> cache[GX_MAX_GL_ORDER + 1];
  static std::once_flag once[GX_MAX_GL_ORDER + 1];
  std::call_once(once[n], [n] {
    std::vector<double> sw(2 * n);
    for (int k = 1; k <= n; ++k) {
      double x = std::cos(M_PI * (k - 0.25) / (n + 0.5)), dp = 1.0;
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = x;
        for (int j = 2; j <= n; ++j) {
          const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
          p0 = p1;
          p1 = p2;
        }
        if (n == 1) p0 = 1.0;
        dp = n * (x * p1 - p0) / ((x - 1.0) * (x + 1.0));
        const double dx = p1 / dp;
        x -= dx;
        if (std::fabs(dx) < 1e-16) {
          // refresh the derivative at the converged root
          p0 = 1.0;
          p1 = x;
          for (int j = 2; j <= n; ++j) {
            const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
            p0 = p1;
            p1 = p2;
          }
          if (n == 1) p0 = 1.0;
          dp = n * (x * p1 - p0) / ((x - 1.0) * (x + 1.0));
          break;
        }
      }
      const double w = 2.0 / ((1.0 - x) * (1.0 + x) * dp * dp);
      sw[2 * (k - 1)] = 0.5 * (x + 1.0);
      sw[2 * (k - 1) + 1] = 0.5 * w;
    }
    cache[n] = std::move(sw);
  });
  return cache[n];
}

template <int O>
J<O> r_spherical(const J<O>& x, const J<O>& y, const J<O>& z) {
  return jsqrt(sq(x) + sq(y) + sq(z) + kTiny);
}

template <int O>
J<O> phi_model(const gx_component& c, const J<O>& x, const J<O>& y,
               const J<O>& z, double t) {
  const double* p = c.p;
  switch (c.kind) {
    case GX_KEPLER: {  // -G m / r
      J<O> r = r_spherical(x, y, z);
      return (-p[0]) / r;
    }
    case GX_HERNQUIST: {  // -G m / (r + r_s)
      J<O> r = r_spherical(x, y, z);
      return (-p[0]) / (r + p[1]);
    }
    case GX_MIYAMOTO_NAGAI: {  // -G m / sqrt(R^2 + (sqrt(z^2+b^2) + a)^2)
      J<O> R2 = sq(x) + sq(y);
      J<O> zp2 = sq(jsqrt(sq(z) + p[2] * p[2]) + p[1]);
      return (-p[0]) / jsqrt(R2 + zp2);
    }
    case GX_NFW: {  // -(G m / r_s) log1p(r/r_s) / (r/r_s)
      J<O> r = r_spherical(x, y, z);
      J<O> s = r / p[1];
      double phi0 = -p[0] / p[1];
      return phi0 * jlog1p(s) / s;
    }
    case GX_PLUMMER: {  // -G m / sqrt(r^2 + r_s^2)
      J<O> r = r_spherical(x, y, z);
      return (-p[0]) / jsqrt(sq(r) + p[1] * p[1]);
    }
    case GX_MONARI_BAR: {
      // prefactor * U(r/Rb) * (R^2/r^2) * cos(2 (phi - phi_b - Omega t))
      double alpha = p[0], R0 = p[1], v0 = p[2], Rb = p[3], phib = p[4],
             Om = p[5];
      J<O> R2 = sq(x) + sq(y);
      J<O> r2 = R2 + sq(z);
      double q = R0 / Rb;
      double pref = alpha * (v0 * v0 / 3.0) * (q * q * q);
      J<O> s = jsqrt(r2) / Rb;
      J<O> U = (s.v >= 1.0) ? -(1.0 / cube(s)) : cube(s) - 2.0;
      J<O> ph = jatan2(y, x);
      J<O> gam = 2.0 * (ph - phib - Om * t);
      return pref * U * (R2 / r2) * jcos(gam);
    }
    case GX_ISOCHRONE: {  // -G m / (r_s + sqrt(r^2 + r_s^2))
      J<O> r = r_spherical(x, y, z);
      return (-p[0]) / (p[1] + jsqrt(sq(r) + p[1] * p[1]));
    }
    case GX_JAFFE: {  // -(G m / r_s) log(1 + r_s / r)
      J<O> r = r_spherical(x, y, z);
      return (-p[0] / p[1]) * jlog1p(p[1] / r);
    }
    case GX_LOGARITHMIC: {  // v_c^2 / 2 log(r_s^2 + r^2)
      J<O> r = r_spherical(x, y, z);
      return (0.5 * p[0] * p[0]) * jlog(p[1] * p[1] + sq(r));
    }
    case GX_LMJ09_LOG: {  // v_c^2 / 2 log(r_s^2 + (x'/q1)^2 + (y'/q2)^2 + (z/q3)^2)
      const double c = std::cos(p[5]), s = std::sin(p[5]);
      J<O> xp = x * c + y * s;
      J<O> yp = -(x * s) + y * c;
      J<O> r2 = sq(xp / p[2]) + sq(yp / p[3]) + sq(z / p[4]);
      return (0.5 * p[0] * p[0]) * jlog(p[1] * p[1] + r2);
    }
    case GX_LONGMURALI_BAR: {  // G m / (2a) log((x' - a + T-) / (x' + a + T+))
      const double a = p[1], b = p[2], cc = p[3];
      const double ca = std::cos(p[4]), sa = std::sin(p[4]);
      J<O> xp = x * ca + y * sa;
      J<O> yp = -(x * sa) + y * ca;
      J<O> yz2 = sq(yp) + sq(b + jsqrt(sq(z) + cc * cc));
      J<O> Tp = jsqrt(sq(a + xp) + yz2);
      J<O> Tm = jsqrt(sq(a - xp) + yz2);
      return (p[0] / (2.0 * a)) * jlog((xp - a + Tm) / (xp + a + Tp));
    }
    case GX_POWERLAW_CUTOFF: {
      // -G m [P(3/2 - alpha/2, s^2) / r + Gamma(1 - alpha/2) / Gamma(3/2 - alpha/2)
      //        Q(1 - alpha/2, s^2) / r_c],  s = r / r_c  (the reference's form
      // with Gamma(5/2 - alpha/2) = (3/2 - alpha/2) Gamma(3/2 - alpha/2) and
      // Phi(inf) = 0 folded in)
      const double a1 = 1.5 - 0.5 * p[1], a0 = 1.0 - 0.5 * p[1], rc = p[2];
      J<O> r = r_spherical(x, y, z);
      J<O> s2 = sq(r / rc);
      const double k = std::exp(std::lgamma(a0) - std::lgamma(a1)) / rc;
      return (-p[0]) * (jgammap(a1, s2) / r + k * (1.0 - jgammap(a0, s2)));
    }
    case GX_BURKERT: {
      // -G m / (r_s C) [pi - 2 (1 + 1/x) atan x + 2 (1 + 1/x) log1p x
      //                 - (1 - 1/x) log1p x^2], x = r / r_s, C = 3 log 2 - pi/2
      const double C = 3.0 * std::log(2.0) - M_PI / 2.0;
      J<O> X = r_spherical(x, y, z) / p[1];
      J<O> xi = 1.0 / X;
      J<O> br = M_PI - 2.0 * (1.0 + xi) * jatan(X) + 2.0 * (1.0 + xi) * jlog1p(X) -
                (1.0 - xi) * jlog1p(sq(X));
      return (-p[0] / (p[1] * C)) * br;
    }
    case GX_KUZMIN: {  // -G m / sqrt(R^2 + (|z| + r_s)^2)
      return (-p[0]) / jsqrt(sq(x) + sq(y) + sq(jabs(z) + p[1]));
    }
    case GX_SATOH: {  // -G m / sqrt(R^2 + z^2 + a (a + 2 sqrt(z^2 + b^2)))
      const double a = p[1], b = p[2];
      return (-p[0]) / jsqrt(sq(x) + sq(y) + sq(z) + a * (a + 2.0 * jsqrt(sq(z) + b * b)));
    }
    case GX_STONE: {
      // -2 G m / (pi (r_h - r_c)) [(r_h atan(r/r_h) - r_c atan(r/r_c)) / r
      //                            + log((r^2 + r_h^2) / (r^2 + r_c^2)) / 2]
      const double rc = p[1], rh = p[2];
      J<O> r = r_spherical(x, y, z);
      J<O> at = (rh * jatan(r / rh) - rc * jatan(r / rc)) / r;
      J<O> lg = 0.5 * jlog((sq(r) + rh * rh) / (sq(r) + rc * rc));
      return (-2.0 * p[0] / (M_PI * (rh - rc))) * (at + lg);
    }
    case GX_HARMONIC_OSCILLATOR: {  // omega^2 r^2 / 2 (example.py:75-85)
      return (0.5 * p[0] * p[0]) * (sq(x) + sq(y) + sq(z));
    }
    case GX_HENON_HEILES: {  // (R^2/2 + c (x^2 y - y^3/3)) / t_s^2 (example.py:159-176)
      const double c = p[0], its2 = 1.0 / (p[1] * p[1]);
      return its2 * (0.5 * (sq(x) + sq(y)) + c * (sq(x) * y - (1.0 / 3.0) * (sq(y) * y)));
    }
    case GX_HARDCUTOFF_NFW: {
      // r <= r_t: Phi_NFW + G m / (r_s + r_t); r > r_t: -G m mu(r_t/r_s) / r,
      // mu(x) = log(1 + x) - x / (1 + x)  (nfw/truncated.py:255-330)
      const double rs = p[1], rt = p[2];
      J<O> r = r_spherical(x, y, z);
      if (r.v <= rt) return (-p[0] / rs) * jlog1p(r / rs) / (r / rs) + p[0] / (rs + rt);
      const double xt = rt / rs;
      return (-p[0] * (std::log1p(xt) - xt / (1.0 + xt))) / r;
    }
    case GX_GNFW: {
      // -G m (B(x)/r + (1 - z^(2-gamma)) / ((2 - gamma) r_s)), z = x/(1 + x),
      // B(x) = B_z(3 - gamma, 0)  (nfw/generalized.py:30-330)
      const double rs = p[1], g = p[2], a = 3.0 - g, b = 2.0 - g;
      J<O> r = r_spherical(x, y, z);
      J<O> X = r / rs;
      J<O> zz = X / (1.0 + X);
      return (-p[0]) * (jgnfw_mass(a, X) / r + (1.0 / (b * rs)) * (1.0 - jpow(zz, b)));
    }
    case GX_TRIAXIAL_GAUSSIAN: {
      // -2 pi G rho0 r_s^2 q1 q2 int_0^1 2 exp(-xi^2/2) / sqrt(A1 A2) ds,
      // r_s^2 xi^2 = s^2 (x^2 + y^2/A1 + z^2/A2), A_i = 1 + (q_i^2 - 1) s^2,
      // rho0 = m / (q1 q2 (2 pi)^(3/2) r_s^3)  (gaussian/triaxial.py:86-221)
      const double rs = p[1], q1sq = p[2] * p[2], q2sq = p[3] * p[3];
      const std::vector<double>& sw = gauss_legendre01((int)p[4]);
      J<O> sum = jconst<O>(0.0);
      for (int k = 0; k < (int)p[4]; ++k) {
        const double s2 = sw[2 * k] * sw[2 * k];
        const double A1 = 1.0 + (q1sq - 1.0) * s2, A2 = 1.0 + (q2sq - 1.0) * s2;
        J<O> xi2 = (s2 / (rs * rs)) * (sq(x) + sq(y) / A1 + sq(z) / A2);
        sum = sum + (sw[2 * k + 1] / std::sqrt(A1 * A2)) * (2.0 * jexp(-0.5 * xi2));
      }
      // 2 pi G rho0 r_s^2 q1 q2 = G m / (sqrt(2 pi) r_s)
      return (-p[0] / (std::sqrt(2.0 * M_PI) * rs)) * sum;
    }
    case GX_GAUSSIAN: {
      // -G m erf(r / (sqrt2 r_s)) / r  (gaussian/base.py:76-95,240-259); for
      // r < sqrt2 r_s through the series of erf(X)/X in X^2, which keeps the
      // derivatives free of cancellation near the centre
      const double rs = p[1], k = 1.0 / (std::sqrt(2.0) * rs);
      J<O> r = r_spherical(x, y, z);
      J<O> X = k * r;
      if (X.v < 1.0) return (-p[0] * k) * jerf_over_sqrt(sq(X));
      return (-p[0]) * jerf(X) / r;
    }
    case GX_TRIAXIAL_HERNQUIST: {
      // -G m / (r' + r_s), r' = sqrt(x^2 + (y/q1)^2 + (z/q2)^2 + tiny)
      // (builtin/hernquist.py:163-182)
      J<O> rp = jsqrt(sq(x) + sq(y / p[2]) + sq(z / p[3]) + kTiny);
      return (-p[0]) / (rp + p[1]);
    }
    case GX_TRIAXIAL_NFW: {
      // -2 pi G rho0 r_s^2 q1 q2 sum_k w_k [2 / (1 + xi_k)] /
      //   sqrt(((q1^2 - 1) s_k^2 + 1)((q2^2 - 1) s_k^2 + 1)),
      // r_s^2 xi_k^2 = s_k^2 (x^2 + y^2 / (1 + (q1^2 - 1) s_k^2)
      //                       + z^2 / (1 + (q2^2 - 1) s_k^2)),
      // rho0 = m / (4 pi r_s^3)  (builtin/nfw/triaxial.py:103-209)
      const double rs = p[1], q1sq = p[2] * p[2], q2sq = p[3] * p[3];
      const std::vector<double>& sw = gauss_legendre01((int)p[4]);
      J<O> sum = jconst<O>(0.0);
      for (int k = 0; k < (int)p[4]; ++k) {
        const double s2 = sw[2 * k] * sw[2 * k];
        const double A1 = 1.0 + (q1sq - 1.0) * s2, A2 = 1.0 + (q2sq - 1.0) * s2;
        J<O> xi = jsqrt(s2 * (sq(x) + sq(y) / A1 + sq(z) / A2)) / rs;
        sum = sum + (sw[2 * k + 1] / std::sqrt(A1 * A2)) * (2.0 / (1.0 + xi));
      }
      // 2 pi G rho0 r_s^2 = G m / (2 r_s)
      return (-p[0] / (2.0 * rs) * p[2] * p[3]) * sum;
    }
    case GX_LEESUTO_NFW: {
      // phi0 [F1 + (e_b^2 + e_c^2)/2 F2 + (e_b^2 sin^2(th) sin^2(ph)
      //   + e_c^2 cos^2(th))/2 F3] (builtin/nfw/leesuto.py:124-164)
      const double rs = p[1], a1 = p[2], a2 = p[3], a3 = p[4];
      const double LOG2 = std::log(2.0);
      const double vc2 = p[0] / rs;
      const double eb2 = 1.0 - (a2 / a1) * (a2 / a1), ec2 = 1.0 - (a3 / a1) * (a3 / a1);
      const double phi0 = vc2 / (LOG2 - 0.5 + (LOG2 - 0.75) * (eb2 + ec2));
      J<O> r = r_spherical(x, y, z);
      J<O> s = r / rs;
      if (s.v == 0.0) return jconst<O>(phi0);
      J<O> L = jlog1p(s);
      J<O> u2 = sq(s);
      J<O> um3 = 1.0 / cube(s);
      J<O> costh2 = sq(z) / sq(r);
      J<O> sinsin2 = sq(y) / sq(r);
      J<O> F1 = -(L / s);
      J<O> F2 = -1.0 / 3.0 + (2.0 * u2 - 3.0 * s + 6.0) / (6.0 * u2) + (1.0 / s - um3) * L;
      J<O> F3 = (u2 - 3.0 * s - 6.0) / (2.0 * u2 * (1.0 + s)) + 3.0 * um3 * L;
      return phi0 * (F1 + (0.5 * (eb2 + ec2)) * F2 + 0.5 * (eb2 * sinsin2 + ec2 * costh2) * F3);
    }
    case GX_VOGELSBERGER_NFW: {
      // -G m log1p(r~ / r_s) / r~, r~ = (r_a + r) r_e / (r_a + r_e),
      // r_e^2 = x^2 + y^2/q1^2 + z^2/q2^2, q2^2 = 3 - q1^2, r_a = a_r r_s
      // (builtin/nfw/vogelsburger08.py:60-110)
      const double rs = p[1], q1sq = p[2] * p[2], q2sq = 3.0 - q1sq, ra = p[3] * rs;
      J<O> re = jsqrt(sq(x) + sq(y) / q1sq + sq(z) / q2sq + kTiny);
      J<O> r = r_spherical(x, y, z);
      J<O> rt = (ra + r) * re / (ra + re);
      return (-p[0]) * jlog1p(rt / rs) / rt;
    }
    default:
      return jconst<O>(0.0);
  }
}

struct Track {
  int64_t n = 0;
  const double* t = nullptr;
  const double* q = nullptr;
  const double* v = nullptr;
  // a dense solution instead (rec != nullptr): n step records of
  // GX_DENSE_RECORD doubles in integration order, written by `method`
  const double* rec = nullptr;
  int method = 0;
};

void dense_track_pos(const Track& tr, double t, double out[3]);

// Cubic Hermite interpolant of a tabulated trajectory (end knots held), or
// a dense solution's own dense output.
void track_pos(const Track& tr, double t, double out[3]) {
  if (tr.rec) {
    dense_track_pos(tr, t, out);
    return;
  }
  if (tr.n <= 0) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  if (tr.n == 1 || t <= tr.t[0]) {
    for (int k = 0; k < 3; ++k) out[k] = tr.q[k];
    return;
  }
  if (t >= tr.t[tr.n - 1]) {
    for (int k = 0; k < 3; ++k) out[k] = tr.q[3 * (tr.n - 1) + k];
    return;
  }
  // largest i with t_i <= t
  int64_t lo = 0, hi = tr.n - 1;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (tr.t[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  double h = tr.t[lo + 1] - tr.t[lo];
  double u = (t - tr.t[lo]) / h;
  double u2 = u * u, u3 = u2 * u;
  double h00 = 2 * u3 - 3 * u2 + 1, h10 = u3 - 2 * u2 + u;
  double h01 = -2 * u3 + 3 * u2, h11 = u3 - u2;
  for (int k = 0; k < 3; ++k)
    out[k] = h00 * tr.q[3 * lo + k] + h10 * h * tr.v[3 * lo + k] +
             h01 * tr.q[3 * (lo + 1) + k] + h11 * h * tr.v[3 * (lo + 1) + k];
}

struct Pot {
  std::vector<gx_component> comps;
  std::vector<Track> tracks;
};

void frame_offset(const Pot& P, const gx_component& c, double t,
                  double d[3]) {
  d[0] = d[1] = d[2] = 0;
  switch (c.frame) {
    case GX_FRAME_OFFSET:
      d[0] = c.f[0];
      d[1] = c.f[1];
      d[2] = c.f[2];
      break;
    case GX_FRAME_ROTATING: {
      double a = c.f[3] * t, ca = std::cos(a), sa = std::sin(a);
      d[0] = c.f[0] * ca - c.f[1] * sa;
      d[1] = c.f[0] * sa + c.f[1] * ca;
      d[2] = c.f[2];
      break;
    }
    case GX_FRAME_TRACK:
      if (c.track >= 0 && c.track < (int)P.tracks.size())
        track_pos(P.tracks[c.track], t, d);
      break;
    default:
      break;
  }
}

template <int O>
J<O> phi_component(const Pot& P, const gx_component& c, const double xyz[3],
                   double t) {
  double d[3];
  frame_offset(P, c, t, d);
  J<O> x = jvar<O>(xyz[0] - d[0], 0);
  J<O> y = jvar<O>(xyz[1] - d[1], 1);
  J<O> z = jvar<O>(xyz[2] - d[2], 2);
  return phi_model<O>(c, x, y, z, t);
}

double potential(const Pot& P, const double xyz[3], double t) {
  double s = 0;
  for (const auto& c : P.comps) s += phi_component<1>(P, c, xyz, t).v;
  return s;
}

void gradient(const Pot& P, const double xyz[3], double t, double g[3]) {
  g[0] = g[1] = g[2] = 0;
  for (const auto& c : P.comps) {
    J<1> j = phi_component<1>(P, c, xyz, t);
    for (int k = 0; k < 3; ++k) g[k] += j.g[k];
  }
}

void hessian(const Pot& P, const double xyz[3], double t, double H[9]) {
  for (int k = 0; k < 9; ++k) H[k] = 0;
  for (const auto& c : P.comps) {
    J<2> j = phi_component<2>(P, c, xyz, t);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) H[3 * a + b] += j.h[a][b];
  }
}

// ------------------------------------------------------------- tableaux
struct Tableau {
  int s = 0;      // stages
  int order = 0;  // order used by the controller (diffrax error_order)
  double c[13] = {};
  double a[13][13] = {};
  double b[13] = {};
  double be[13] = {};  // b_sol - b_embedded
  // Dense output over s + 1 increments: k_1..k_s and the first-same-as-last
  // increment k_{s+1} = h f(t1, y1) (Dopri5's own last stage is already at
  // y1, so for it k_{s+1} equals k_s up to rounding).
  int dense = 0;        // kDensePoly or kDenseQuartic
  int dn = 0;           // polynomial degree of the weights in theta
  double dw[14][6] = {};  // kDensePoly: b_i(th) = th * polyval(dw[i], th),
                          // coefficients from th^(dn-1) down to th^0
  double cmid[7] = {};    // kDenseQuartic: y(1/2) = y0 + sum cmid_i k_i
};

constexpr int kDensePoly = 0;
constexpr int kDenseQuartic = 1;

Tableau make_dopri8() {
  // Prince & Dormand (1981), RK8(7)13M.
  Tableau T;
  T.s = 13;
  T.order = 8;
  const double c[13] = {0.0,
                        1.0 / 18,
                        1.0 / 12,
                        1.0 / 8,
                        5.0 / 16,
                        3.0 / 8,
                        59.0 / 400,
                        93.0 / 200,
                        5490023248.0 / 9719169821.0,
                        13.0 / 20,
                        1201146811.0 / 1299019798.0,
                        1.0,
                        1.0};
  for (int i = 0; i < 13; ++i) T.c[i] = c[i];
  auto& a = T.a;
  a[1][0] = 1.0 / 18;
  a[2][0] = 1.0 / 48;
  a[2][1] = 1.0 / 16;
  a[3][0] = 1.0 / 32;
  a[3][2] = 3.0 / 32;
  a[4][0] = 5.0 / 16;
  a[4][2] = -75.0 / 64;
  a[4][3] = 75.0 / 64;
  a[5][0] = 3.0 / 80;
  a[5][3] = 3.0 / 16;
  a[5][4] = 3.0 / 20;
  a[6][0] = 29443841.0 / 614563906.0;
  a[6][3] = 77736538.0 / 692538347.0;
  a[6][4] = -28693883.0 / 1125000000.0;
  a[6][5] = 23124283.0 / 1800000000.0;
  a[7][0] = 16016141.0 / 946692911.0;
  a[7][3] = 61564180.0 / 158732637.0;
  a[7][4] = 22789713.0 / 633445777.0;
  a[7][5] = 545815736.0 / 2771057229.0;
  a[7][6] = -180193667.0 / 1043307555.0;
  a[8][0] = 39632708.0 / 573591083.0;
  a[8][3] = -433636366.0 / 683701615.0;
  a[8][4] = -421739975.0 / 2616292301.0;
  a[8][5] = 100302831.0 / 723423059.0;
  a[8][6] = 790204164.0 / 839813087.0;
  a[8][7] = 800635310.0 / 3783071287.0;
  a[9][0] = 246121993.0 / 1340847787.0;
  a[9][3] = -37695042795.0 / 15268766246.0;
  a[9][4] = -309121744.0 / 1061227803.0;
  a[9][5] = -12992083.0 / 490766935.0;
  a[9][6] = 6005943493.0 / 2108947869.0;
  a[9][7] = 393006217.0 / 1396673457.0;
  a[9][8] = 123872331.0 / 1001029789.0;
  a[10][0] = -1028468189.0 / 846180014.0;
  a[10][3] = 8478235783.0 / 508512852.0;
  a[10][4] = 1311729495.0 / 1432422823.0;
  a[10][5] = -10304129995.0 / 1701304382.0;
  a[10][6] = -48777925059.0 / 3047939560.0;
  a[10][7] = 15336726248.0 / 1032824649.0;
  a[10][8] = -45442868181.0 / 3398467696.0;
  a[10][9] = 3065993473.0 / 597172653.0;
  a[11][0] = 185892177.0 / 718116043.0;
  a[11][3] = -3185094517.0 / 667107341.0;
  a[11][4] = -477755414.0 / 1098053517.0;
  a[11][5] = -703635378.0 / 230739211.0;
  a[11][6] = 5731566787.0 / 1027545527.0;
  a[11][7] = 5232866602.0 / 850066563.0;
  a[11][8] = -4093664535.0 / 808688257.0;
  a[11][9] = 3962137247.0 / 1805957418.0;
  a[11][10] = 65686358.0 / 487910083.0;
  a[12][0] = 403863854.0 / 491063109.0;
  a[12][3] = -5068492393.0 / 434740067.0;
  a[12][4] = -411421997.0 / 543043805.0;
  a[12][5] = 652783627.0 / 914296604.0;
  a[12][6] = 11173962825.0 / 925320556.0;
  a[12][7] = -13158990841.0 / 6184727034.0;
  a[12][8] = 3936647629.0 / 1978049680.0;
  a[12][9] = -160528059.0 / 685178525.0;
  a[12][10] = 248638103.0 / 1413531060.0;
  const double b[13] = {14005451.0 / 335480064.0,
                        0,
                        0,
                        0,
                        0,
                        -59238493.0 / 1068277825.0,
                        181606767.0 / 758867731.0,
                        561292985.0 / 797845732.0,
                        -1041891430.0 / 1371343529.0,
                        760417239.0 / 1151165299.0,
                        118820643.0 / 751138087.0,
                        -528747749.0 / 2220607170.0,
                        1.0 / 4};
  const double bh[13] = {13451932.0 / 455176623.0,
                         0,
                         0,
                         0,
                         0,
                         -808719846.0 / 976000145.0,
                         1757004468.0 / 5645159321.0,
                         656045339.0 / 265891186.0,
                         -3867574721.0 / 1518517206.0,
                         465885868.0 / 322736535.0,
                         53011238.0 / 667516719.0,
                         2.0 / 45,
                         0};
  for (int i = 0; i < 13; ++i) {
    T.b[i] = b[i];
    T.be[i] = b[i] - bh[i];
  }
  // diffrax 0.7.0's Dopri8 interpolant (_solver/dopri8.py,
  // _Dopri8Interpolation; third-party, not vendored; used through
  // SaveAt(ts) at legacy/integrator.py:227-230 and orbit/solver.py:794):
  // y(t0 + th h) = y0 + sum_{i=1}^{14} b_i(th) k_i with b_i(th) =
  // th * (quintic in th), over the 13 stages and the first-same-as-last
  // increment.  Stages 2-5 carry no weight.  Row 1 is not tabulated: the
  // weights sum to th (the interpolant reproduces y' = const exactly), which
  // fixes it from the others.  The reference's Kepler integrate_field
  // doctest (dynamics/_src/solver.py:342-365) pins the off-grid values.
  T.dense = kDensePoly;
  T.dn = 6;
  const double w[9][6] = {
      {-39.6107919852202505218, 116.4422149550342161651, -121.4999627731334642623,
       52.2273532792945524050, -7.6142658045872677172, 0},
      {20.3761213808791436958, -67.1451318825957197185, 83.1721004639847717481,
       -46.8919164181093621583, 10.7281392630428866124, 0},
      {7.3347098826795362023, -16.5672243527496524646, 9.5724507555993664382,
       -0.1890893225010595467, 0.5526637063753648783, 0},
      {32.8801774352459155182, -89.9916014847245016028, 87.8406057677205645007,
       -35.7075975946222072821, 4.2186562625665153803, 0},
      {-10.1588990526426760954, 22.6237489648532849093, -17.4152107770762969005,
       6.2736448083240352160, -0.6627209125361597559, 0},
      {-12.5401268098782561200, 32.2362340167355370113, -28.5903289514790976966,
       10.3160881272450748458, -1.2636789001135462218, 0},
      {29.5553001484516038033, -82.1020315488359848644, 81.6630950584341412934,
       -34.7650769866611817349, 5.4106037898590422230, 0},
      {-41.7923486424390588923, 116.2662185791119533462, -114.9375291377009418170,
       47.7457971078225540396, -7.0321379067945741781, 0},
      {20.3006925822100825485, -53.9020777466385396792, 50.2558364226176017553,
       -19.0082099341608028453, 2.3537586759714983486, 0}};
  // rows of stages 6..14 (indices 5..13)
  for (int r = 0; r < 9; ++r)
    for (int k = 0; k < 6; ++k) T.dw[5 + r][k] = w[r][k];
  for (int k = 0; k < 6; ++k) {
    double acc = (k == 5) ? 1.0 : 0.0;
    for (int i = 5; i < 14; ++i) acc -= T.dw[i][k];
    T.dw[0][k] = acc;
  }
  return T;
}

Tableau make_dopri5() {
  // Dormand & Prince (1980), RK5(4)7M, with the embedded weights diffrax
  // uses for its error estimate.
  Tableau T;
  T.s = 7;
  T.order = 5;
  const double c[7] = {0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
  for (int i = 0; i < 7; ++i) T.c[i] = c[i];
  auto& a = T.a;
  a[1][0] = 1.0 / 5;
  a[2][0] = 3.0 / 40;
  a[2][1] = 9.0 / 40;
  a[3][0] = 44.0 / 45;
  a[3][1] = -56.0 / 15;
  a[3][2] = 32.0 / 9;
  a[4][0] = 19372.0 / 6561;
  a[4][1] = -25360.0 / 2187;
  a[4][2] = 64448.0 / 6561;
  a[4][3] = -212.0 / 729;
  a[5][0] = 9017.0 / 3168;
  a[5][1] = -355.0 / 33;
  a[5][2] = 46732.0 / 5247;
  a[5][3] = 49.0 / 176;
  a[5][4] = -5103.0 / 18656;
  a[6][0] = 35.0 / 384;
  a[6][2] = 500.0 / 1113;
  a[6][3] = 125.0 / 192;
  a[6][4] = -2187.0 / 6784;
  a[6][5] = 11.0 / 84;
  const double b[7] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192,
                       -2187.0 / 6784, 11.0 / 84, 0};
  const double be[7] = {35.0 / 384 - 1951.0 / 21600,
                        0,
                        500.0 / 1113 - 22642.0 / 50085,
                        125.0 / 192 - 451.0 / 720,
                        -2187.0 / 6784 + 12231.0 / 42400,
                        11.0 / 84 - 649.0 / 6300,
                        -1.0 / 60.0};
  for (int i = 0; i < 7; ++i) {
    T.b[i] = b[i];
    T.be[i] = be[i];
  }
  // diffrax 0.7.0's Dopri5 interpolant (_solver/dopri5.py,
  // _Dopri5Interpolation, a FourthOrderPolynomialInterpolation from
  // _local_interpolation.py; third-party): the quartic through y0, y1, the
  // end-point slopes k_1, k_7 and Shampine's (1986) midpoint value
  // y0 + sum cmid_i k_i.
  T.dense = kDenseQuartic;
  T.dn = 4;
  const double cm[7] = {6025192743.0 / 30085553152.0 / 2,
                        0,
                        51252292925.0 / 65400821598.0 / 2,
                        -2691868925.0 / 45128329728.0 / 2,
                        187940372067.0 / 1594534317056.0 / 2,
                        -1776094331.0 / 19743644256.0 / 2,
                        11237099.0 / 235043384.0 / 2};
  for (int i = 0; i < 7; ++i) T.cmid[i] = cm[i];
  return T;
}

const Tableau& tableau(int method) {
  static const Tableau d8 = make_dopri8();
  static const Tableau d5 = make_dopri5();
  return method == GX_DOPRI5 ? d5 : d8;
}

// -------------------------------------------------------------- stepping
// Vector field of Hamilton's equations in the solver's time variable
// tau = dir * t: dy/dtau = dir * (p, -grad Phi(q, dir*tau)).
//
// NBodyField (dynamics/_src/orbit/field_nbody.py:94-131) couples B bodies:
// a_j = sum_{i != j} G m_i (x_i - x_j) / (|x_i - x_j|^2 + eps^2)^(3/2)
//       - grad Phi_ext(x_j, t).
struct NBody {
  std::vector<double> gm;  // G m_i
  double eps2 = 0.0;
};

struct Field {
  const Pot* P;
  double dir;
  const NBody* nb = nullptr;
  void operator()(double tau, const double* y, double* f) const {
    double g[3];
    gradient(*P, y, dir * tau, g);
    f[0] = dir * y[3];
    f[1] = dir * y[4];
    f[2] = dir * y[5];
    f[3] = -dir * g[0];
    f[4] = -dir * g[1];
    f[5] = -dir * g[2];
  }
  // the field of all B states y[6B] at once
  void all(double tau, const double* y, int64_t B, double* f) const {
    if (!nb) {
      for (int64_t o = 0; o < B; ++o) (*this)(tau, y + 6 * o, f + 6 * o);
      return;
    }
    for (int64_t j = 0; j < B; ++j) {
      const double* yj = y + 6 * j;
      double g[3];
      gradient(*P, yj, dir * tau, g);
      double a[3] = {-g[0], -g[1], -g[2]};
      for (int64_t i = 0; i < B; ++i) {
        if (i == j) continue;
        const double* yi = y + 6 * i;
        const double d[3] = {yi[0] - yj[0], yi[1] - yj[1], yi[2] - yj[2]};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + nb->eps2;
        const double w = nb->gm[i] / (r2 * std::sqrt(r2));
        for (int m = 0; m < 3; ++m) a[m] += w * d[m];
      }
      double* fj = f + 6 * j;
      for (int m = 0; m < 3; ++m) {
        fj[m] = dir * yj[3 + m];
        fj[3 + m] = dir * a[m];
      }
    }
  }
};

// One explicit RK step of size h from (tau, y[6*B]) for B orbits sharing the
// step: stage increments k = h * f, Y_i = y + sum_j a_ij k_j,
// y1 = y + sum_i b_i k_i, err = sum_i be_i k_i.
// kall, if given, receives the increments k_1..k_s ([s][B][6]).
void rk_step(const Tableau& T, const Field& F, double tau, const double* y,
             double h, int64_t B, double* y1, double* err,
             std::vector<double>* kall = nullptr) {
  const int s = T.s;
  std::vector<double> k((size_t)s * 6 * B), Y(6 * B), fb(6 * B);
  for (int i = 0; i < s; ++i) {
    double ti = tau + T.c[i] * h;
    for (int64_t o = 0; o < B; ++o) {
      for (int m = 0; m < 6; ++m) {
        double acc = 0;
        for (int j = 0; j < i; ++j)
          if (T.a[i][j] != 0.0) acc += T.a[i][j] * k[((size_t)j * B + o) * 6 + m];
        Y[o * 6 + m] = y[o * 6 + m] + acc;
      }
    }
    F.all(ti, Y.data(), B, fb.data());
    for (int64_t o = 0; o < B; ++o)
      for (int m = 0; m < 6; ++m) k[((size_t)i * B + o) * 6 + m] = fb[o * 6 + m] * h;
  }
  for (int64_t o = 0; o < B; ++o)
    for (int m = 0; m < 6; ++m) {
      double sb = 0, se = 0;
      for (int i = 0; i < s; ++i) {
        double ki = k[((size_t)i * B + o) * 6 + m];
        if (T.b[i] != 0.0) sb += T.b[i] * ki;
        if (T.be[i] != 0.0) se += T.be[i] * ki;
      }
      y1[o * 6 + m] = y[o * 6 + m] + sb;
      if (err) err[o * 6 + m] = se;
    }
  if (kall) kall->swap(k);
}

// Number of increments the dense output reads: Dopri8 adds the
// first-same-as-last increment h f(t1, y1) as a 14th.
int dense_slots(const Tableau& T) {
  return T.dense == kDensePoly ? T.s + 1 : T.s;
}

// The step's dense output at th in [0, 1]: y(tau + th h) for B states,
// from y0, y1 and the increments k[dense_slots][B][6].
void dense_eval(const Tableau& T, const double* y0, const double* y1,
                const double* k, int64_t B, double th, double* out) {
  const int ns = dense_slots(T);
  const size_t n6 = 6 * (size_t)B;
  if (T.dense == kDensePoly) {
    double w[14];
    for (int i = 0; i < ns; ++i) {
      double acc = 0.0;
      for (int d = 0; d < T.dn; ++d) acc = acc * th + T.dw[i][d];
      w[i] = acc * th;
    }
    for (size_t j = 0; j < n6; ++j) {
      double acc = 0.0;
      for (int i = 0; i < ns; ++i) acc += w[i] * k[i * n6 + j];
      out[j] = y0[j] + acc;
    }
    return;
  }
  // quartic through y0, y1, the end slopes and the midpoint value
  for (size_t j = 0; j < n6; ++j) {
    double ymid = y0[j];
    for (int i = 0; i < T.s; ++i) ymid += T.cmid[i] * k[i * n6 + j];
    const double f0 = k[j], f1 = k[(size_t)(T.s - 1) * n6 + j];
    const double a = 2 * (f1 - f0) - 8 * (y1[j] + y0[j]) + 16 * ymid;
    const double b = 5 * f0 - 3 * f1 + 18 * y0[j] + 14 * y1[j] - 32 * ymid;
    const double c = f1 - 4 * f0 - 11 * y0[j] - 5 * y1[j] + 16 * ymid;
    out[j] = (((a * th + b) * th + c) * th + f0) * th + y0[j];
  }
}

// Position of a dense-solution track at t (galax's Orbit interpolant over
// diffrax's DenseInterpolation, orbit/interp.py:42-145): the record whose
// step holds t, then the method's first-order dense output (dense_eval)
// over it, with the increments rebuilt from the record: stage velocities
// v_i = v0 + h sum_j a_ij g_j, increments k_i = h (v_i, g_i), and for
// Dopri8's extra slot k = h (v1, g_end), v1 = v0 + h sum_j b_j g_j.  Outside
// the integrated interval the end position is held.
void record_pos(const Tableau& T, const double* r, double th, double out[3]) {
  const int s = T.s, ns = dense_slots(T);
  const double h = r[1];
  const double* g = r + 8;
  double y0[6], y1[6];
  double k[14 * 6] = {};
  for (int m = 0; m < 3; ++m) {
    y0[m] = r[2 + m];
    y0[3 + m] = r[5 + m];
  }
  double q1[3] = {y0[0], y0[1], y0[2]}, v1[3] = {y0[3], y0[4], y0[5]};
  for (int i = 0; i < s; ++i) {
    for (int m = 0; m < 3; ++m) {
      double vi = y0[3 + m];
      for (int j = 0; j < i; ++j) vi += h * T.a[i][j] * g[3 * j + m];
      k[(size_t)i * 6 + m] = h * vi;
      k[(size_t)i * 6 + 3 + m] = h * g[3 * i + m];
      q1[m] += T.b[i] * h * vi;
      v1[m] += T.b[i] * h * g[3 * i + m];
    }
  }
  for (int m = 0; m < 3; ++m) {
    y1[m] = q1[m];
    y1[3 + m] = v1[m];
  }
  if (ns > s)
    for (int m = 0; m < 3; ++m) {
      k[(size_t)s * 6 + m] = h * v1[m];
      k[(size_t)s * 6 + 3 + m] = h * g[3 * s + m];
    }
  double o[6];
  dense_eval(T, y0, y1, k, 1, th, o);
  for (int m = 0; m < 3; ++m) out[m] = o[m];
}

void dense_track_pos(const Track& tr, double t, double out[3]) {
  const Tableau& T = tableau(tr.method);
  const int64_t n = tr.n;
  const bool fwd = tr.rec[1] > 0.0;
  // records by increasing time: j -> record idx(j); left end of each
  auto idx = [&](int64_t j) { return fwd ? j : n - 1 - j; };
  auto left = [&](int64_t j) {
    const double* r = tr.rec + GX_DENSE_RECORD * idx(j);
    return fwd ? r[0] : r[0] + r[1];
  };
  const double* first = tr.rec + GX_DENSE_RECORD * idx(0);
  const double* last = tr.rec + GX_DENSE_RECORD * idx(n - 1);
  const double t_lo = left(0), t_hi = fwd ? last[0] + last[1] : last[0];
  if (t <= t_lo) {
    record_pos(T, first, fwd ? 0.0 : 1.0, out);
    return;
  }
  if (t >= t_hi) {
    record_pos(T, last, fwd ? 1.0 : 0.0, out);
    return;
  }
  int64_t lo = 0, hi = n - 1;  // last j with left(j) <= t
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (left(mid) <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const double* r = tr.rec + GX_DENSE_RECORD * idx(lo);
  record_pos(T, r, (t - r[0]) / r[1], out);
}

double rms(const double* x, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
  return std::sqrt(s / (double)n);
}

bool all_finite(const double* x, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(x[i])) return false;
  return true;
}

// diffrax _select_initial_step (Hairer, Norsett & Wanner, II.4: the final
// exponent is 1/(p+1) for a method of order p).
double initial_step(const Field& F, double tau0, const double* y0, int64_t n6,
                    double rtol, double atol, double error_order) {
  std::vector<double> f0(n6), f1(n6), y1(n6), tmp(n6), scale(n6);
  F.all(tau0, y0, n6 / 6, f0.data());
  for (int64_t i = 0; i < n6; ++i) scale[i] = atol + std::fabs(y0[i]) * rtol;
  for (int64_t i = 0; i < n6; ++i) tmp[i] = y0[i] / scale[i];
  double d0 = rms(tmp.data(), n6);
  for (int64_t i = 0; i < n6; ++i) tmp[i] = f0[i] / scale[i];
  double d1 = rms(tmp.data(), n6);
  bool cond = (d0 < 1e-5) || (d1 < 1e-5);
  double d1_ = cond ? 1.0 : d1;
  double h0 = cond ? 1e-6 : 0.01 * (d0 / d1_);
  for (int64_t i = 0; i < n6; ++i) y1[i] = y0[i] + f0[i] * h0;
  F.all(tau0 + h0, y1.data(), n6 / 6, f1.data());
  for (int64_t i = 0; i < n6; ++i) tmp[i] = (f1[i] - f0[i]) / scale[i];
  double d2 = rms(tmp.data(), n6) / h0;
  double maxd = std::max(d1, d2);
  double h1 = (maxd <= 1e-15) ? std::max(1e-6, h0 * 1e-3)
                              : std::pow(0.01 / maxd, 1.0 / (error_order + 1.0));
  return std::min(100 * h0, h1);
}

// diffrax _clip_to_end: avoid a sliver step before t1.
double clip_to_end(double tprev, double tnext, double t1, bool keep) {
  const double tol = 1e-10;
  if (tnext > t1 - tol) return keep ? t1 : tprev + 0.5 * (t1 - tprev);
  return tnext;
}

struct SolveOut {
  int32_t status = GX_SUCCESS;
  int64_t nsteps = 0;
  int64_t naccept = 0;
};

// Solve B orbits that share one step sequence (B == 1 is the independent
// per-orbit solve).  y0[6B]; save times ts[T] in t; out[B][T][6].
SolveOut solve(const Pot& P, const gx_solver_cfg& cfg, const double* y0,
               double t0, double t1, const double* ts, int T, int64_t B,
               double* out, const NBody* nb = nullptr) {
  SolveOut R;
  const Tableau& Tb = tableau(cfg.method);
  const int64_t n6 = 6 * B;
  auto write = [&](int s, const double* y) {
    for (int64_t o = 0; o < B; ++o)
      for (int m = 0; m < 6; ++m)
        out[((size_t)o * T + s) * 6 + m] = y[o * 6 + m];
  };
  std::vector<double> y(y0, y0 + n6), y1(n6), err(n6), ys(n6), kst, f1(n6);
  if (t0 == t1) {
    for (int s = 0; s < T; ++s) write(s, y.data());
    return R;
  }
  const double dir = (t0 < t1) ? 1.0 : -1.0;
  Field F{&P, dir, nb};
  const double tau1 = dir * t1;
  double tau = dir * t0;
  int si = 0;
  // saves at (or before) the start time
  while (si < T && dir * ts[si] <= tau) write(si++, y.data());

  const bool pid = cfg.controller == GX_PID;
  const double dtmax = (cfg.dtmax > 0) ? cfg.dtmax : INFINITY;
  const double dtmin = (cfg.dtmin > 0) ? cfg.dtmin : 0.0;
  double h0;
  bool at_dtmin = false;
  if (cfg.dt0 > 0) {
    h0 = cfg.dt0;
  } else if (pid && dtmin > 0) {
    // With dtmin set and no dt0 the first step is dtmin: the reference's
    // integrate_field doctest (dynamics/_src/solver.py:342-365, default
    // solver :226-232, dtmin 0.05 forced) reaches its printed end point only
    // from that first step; the automatic rule (0.130 Myr there) misses it
    // by 8.5e-6.
    h0 = dtmin;
  } else {
    h0 = initial_step(F, tau, y.data(), n6, cfg.rtol, cfg.atol, Tb.order);
  }
  if (pid) {
    h0 = std::min(h0, dtmax);
    if (dtmin > 0) {
      at_dtmin = h0 <= dtmin;
      h0 = std::max(h0, dtmin);
    }
  }
  double tnext = std::min(tau + h0, tau1);
  double prev_inv = 1.0, prev_prev_inv = 1.0;
  const double order = Tb.order;
  const double c1 = (cfg.icoeff + cfg.pcoeff + cfg.dcoeff) / order;
  const double c2 = -(cfg.pcoeff + 2 * cfg.dcoeff) / order;
  const double c3 = cfg.dcoeff / order;

  // "unbounded" (max_steps <= 0) is a safety cap of 10,000,000 attempts,
  // as in the GPU library
  const int64_t cap =
      (cfg.max_steps > 0 && cfg.max_steps < 10000000) ? cfg.max_steps : 10000000;
  while (tau < tau1) {
    if (R.nsteps >= cap) {
      R.status = GX_MAX_STEPS_REACHED;
      break;
    }
    const double h = tnext - tau;
    rk_step(Tb, F, tau, y.data(), h, B, y1.data(), err.data(), &kst);
    R.nsteps++;
    if (!all_finite(y1.data(), n6)) {
      R.status = GX_NONFINITE;
      break;
    }
    bool keep = true;
    double next_tau1;
    if (pid) {
      std::vector<double> sc(n6);
      for (int64_t i = 0; i < n6; ++i) {
        double ym = std::max(std::fabs(y[i]), std::fabs(y1[i]));
        sc[i] = err[i] / (cfg.atol + ym * cfg.rtol);
      }
      double scaled = rms(sc.data(), n6);
      keep = scaled < 1.0;
      if (dtmin > 0) keep = keep || at_dtmin;
      double inv = 1.0 / scaled;
      double f1 = (c1 == 0) ? 1.0 : std::pow(inv, c1);
      double f2 = (c2 == 0) ? 1.0 : std::pow(prev_inv, c2);
      double f3 = (c3 == 0) ? 1.0 : std::pow(prev_prev_inv, c3);
      double fmin = keep ? 1.0 : cfg.factormin;
      double factor =
          std::min(std::max(cfg.safety * f1 * f2 * f3, fmin), cfg.factormax);
      double dt = h * factor;
      dt = std::min(dt, dtmax);
      if (dtmin > 0) {
        if (!cfg.force_dtmin && dt < dtmin) {
          R.status = GX_DT_MIN_REACHED;
        }
        at_dtmin = dt <= dtmin;
        dt = std::max(dt, dtmin);
      }
      if (keep) {
        prev_prev_inv = prev_inv;
        prev_inv = (inv == 0.0 || std::isinf(inv)) ? 1.0 : inv;
      }
      double nt0 = keep ? tnext : tau;
      next_tau1 = nt0 + dt;
    } else {
      next_tau1 = tnext + cfg.dt0;
    }
    if (keep) {
      // saves in (tau, tnext]: the step end, or the step's dense output
      bool have_fsal = Tb.dense != kDensePoly;
      while (si < T && dir * ts[si] <= tnext) {
        double ts_tau = dir * ts[si];
        if (ts_tau == tnext) {
          write(si, y1.data());
        } else {
          if (!have_fsal) {
            F.all(tnext, y1.data(), B, f1.data());
            kst.resize((size_t)(Tb.s + 1) * n6);
            for (int64_t j = 0; j < n6; ++j) kst[(size_t)Tb.s * n6 + j] = h * f1[j];
            have_fsal = true;
          }
          dense_eval(Tb, y.data(), y1.data(), kst.data(), B, (ts_tau - tau) / h, ys.data());
          write(si, ys.data());
        }
        ++si;
      }
      y.swap(y1);
      R.naccept++;
    }
    double tprev = std::min(keep ? tnext : tau, tau1);
    tnext = clip_to_end(tprev, next_tau1, tau1, keep);
    tau = tprev;
    if (R.status == GX_DT_MIN_REACHED) break;
  }
  // any remaining saves (only if stopped early) get NaN
  if (R.status != GX_SUCCESS) {
    std::vector<double> nanv(n6, NAN);
    while (si < T) write(si++, nanv.data());
  }
  return R;
}

Pot make_pot(const gx_component* comps, int ncomp, const gx_track* tracks,
             int ntracks) {
  Pot P;
  P.comps.assign(comps, comps + ncomp);
  for (int i = 0; i < ntracks; ++i) {
    Track t;
    t.n = tracks[i].n;
    t.t = tracks[i].t;
    t.q = tracks[i].q;
    t.v = tracks[i].v;
    t.rec = tracks[i].rec;
    t.method = tracks[i].method;
    P.tracks.push_back(t);
  }
  return P;
}

}  // namespace

extern "C" {

// what: 0 = Phi, 1 = grad Phi, 2 = Hessian (row-major 3x3), 3 = -grad Phi
int orc_eval(const gx_component* comps, int ncomp, const gx_track* tracks,
             int ntracks, int what, const double* xyz, const double* t,
             int t_stride, int64_t n, double* out) {
  if (!comps && ncomp > 0) return GX_EINVAL;
  Pot P = make_pot(comps, ncomp, tracks, ntracks);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double ti = t[t_stride ? i : 0];
    const double* q = xyz + 3 * i;
    if (what == 0) {
      out[i] = potential(P, q, ti);
    } else if (what == 1 || what == 3) {
      double g[3];
      gradient(P, q, ti, g);
      double s = (what == 3) ? -1.0 : 1.0;
      for (int k = 0; k < 3; ++k) out[3 * i + k] = s * g[k];
    } else {
      hessian(P, q, ti, out + 9 * i);
    }
  }
  return GX_OK;
}

// Tableau export: fills c[s], a[s*s] (row-major), b[s], be[s]; returns s.
// Gauss-Legendre rule on [0, 1] of order n (1..GX_MAX_GL_ORDER): s, w of n.
int orc_gauss_legendre(int n, double* s, double* w) {
  if (n < 1 || n > GX_MAX_GL_ORDER || !s || !w) return GX_EINVAL;
  const std::vector<double>& sw = gauss_legendre01(n);
  for (int k = 0; k < n; ++k) {
    s[k] = sw[2 * k];
    w[k] = sw[2 * k + 1];
  }
  return GX_OK;
}

int orc_tableau(int method, double* c, double* a, double* b, double* be,
                int* order) {
  const Tableau& T = tableau(method);
  for (int i = 0; i < T.s; ++i) {
    c[i] = T.c[i];
    b[i] = T.b[i];
    be[i] = T.be[i];
    for (int j = 0; j < T.s; ++j) a[i * T.s + j] = T.a[i][j];
  }
  *order = T.order;
  return T.s;
}

// The dense output's weights at th: out(th) - y0 = sum_i w_i k_i, one weight
// per increment (dense_slots of them, returned), read off dense_eval with
// one unit increment at a time.
int orc_dense_weights(int method, double th, double* w) {
  const Tableau& T = tableau(method);
  const int ns = dense_slots(T);
  std::vector<double> k((size_t)ns * 6, 0.0), y0(6, 0.0), y1(6), o(6);
  for (int i = 0; i < ns; ++i) {
    std::fill(k.begin(), k.end(), 0.0);
    for (int m = 0; m < 6; ++m) k[(size_t)i * 6 + m] = 1.0;
    const double bi = i < T.s ? T.b[i] : 0.0;
    for (int m = 0; m < 6; ++m) y1[m] = bi;
    dense_eval(T, y0.data(), y1.data(), k.data(), 1, th, o.data());
    w[i] = o[0];
  }
  return ns;
}

// One step of size h from (t, y[6B]) for B orbits (the OrbitSolver.step
// doctest pins this, dynamics/_src/orbit/solver.py:300-321).
int orc_rk_step(const gx_component* comps, int ncomp, const gx_track* tracks,
                int ntracks, int method, double t, const double* y, double h,
                int64_t B, double* y1, double* err) {
  Pot P = make_pot(comps, ncomp, tracks, ntracks);
  Field F{&P, 1.0};
  rk_step(tableau(method), F, t, y, h, B, y1, err);
  return GX_OK;
}

// Integrate n orbits.  shared != 0: one solve with a shared step sequence
// over all n (t0[0], t1[0], shared ts).  Otherwise independent solves,
// parallel over orbits with nthreads OpenMP threads (0 = default).
int orc_integrate(const gx_component* comps, int ncomp, const gx_track* tracks,
                  int ntracks, const gx_solver_cfg* cfg, const double* w0,
                  const double* t0, const double* t1, int t_stride,
                  const double* ts, int T, int ts_mode, int64_t n, int shared,
                  int nthreads, double* out, int32_t* status, int64_t* nsteps,
                  int64_t* naccept) {
  if (!cfg || !w0 || !t0 || !t1 || !out || n < 0 || T < 1) return GX_EINVAL;
  if (cfg->controller == GX_CONSTANT && !(cfg->dt0 > 0)) return GX_EINVAL;
  if (ts_mode == GX_TS_FINAL && T != 1) return GX_EINVAL;
  if ((ts_mode == GX_TS_SHARED || ts_mode == GX_TS_PER_ORBIT) && !ts)
    return GX_EINVAL;
  Pot P = make_pot(comps, ncomp, tracks, ntracks);
  if (shared) {
    if (ts_mode == GX_TS_PER_ORBIT) return GX_EINVAL;
    std::vector<double> tsv(T);
    for (int s = 0; s < T; ++s) {
      if (ts_mode == GX_TS_SHARED)
        tsv[s] = ts[s];
      else if (ts_mode == GX_TS_FINAL)
        tsv[s] = t1[0];
      else
        tsv[s] = (s == T - 1) ? t1[0]
                              : t0[0] + s * ((t1[0] - t0[0]) / (T - 1));
    }
    SolveOut R = solve(P, *cfg, w0, t0[0], t1[0], tsv.data(), T, n, out);
    for (int64_t i = 0; i < n; ++i) {
      if (status) status[i] = R.status;
      if (nsteps) nsteps[i] = R.nsteps;
      if (naccept) naccept[i] = R.naccept;
    }
    return GX_OK;
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> tsv(T);
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < n; ++i) {
      double a = t0[t_stride ? i : 0], b = t1[t_stride ? i : 0];
      for (int s = 0; s < T; ++s) {
        if (ts_mode == GX_TS_SHARED)
          tsv[s] = ts[s];
        else if (ts_mode == GX_TS_PER_ORBIT)
          tsv[s] = ts[(size_t)i * T + s];
        else if (ts_mode == GX_TS_FINAL)
          tsv[s] = b;
        else
          tsv[s] = (s == T - 1) ? b : a + s * ((b - a) / (T - 1));
      }
      SolveOut R =
          solve(P, *cfg, w0 + 6 * i, a, b, tsv.data(), T, 1, out + (size_t)i * T * 6);
      if (status) status[i] = R.status;
      if (nsteps) nsteps[i] = R.nsteps;
      if (naccept) naccept[i] = R.naccept;
    }
  }
  return GX_OK;
}

// N-body system (NBodyField with an external potential): B bodies with
// G m_i = gm[i] and softening eps, integrated as one system (shared steps,
// RMS error over all 6B components) from t0 to t1 with saves at ts[T].
int orc_nbody_integrate(const gx_component* comps, int ncomp,
                        const gx_track* tracks, int ntracks, const double* gm,
                        double eps, const gx_solver_cfg* cfg, const double* w0,
                        double t0, double t1, const double* ts, int T,
                        int64_t B, double* out, int32_t* status,
                        int64_t* nsteps, int64_t* naccept) {
  if (!cfg || !w0 || !gm || !ts || !out || B < 1 || T < 1) return GX_EINVAL;
  if (cfg->controller == GX_CONSTANT && !(cfg->dt0 > 0)) return GX_EINVAL;
  Pot P = make_pot(comps, ncomp, tracks, ntracks);
  NBody nb;
  nb.gm.assign(gm, gm + B);
  nb.eps2 = eps * eps;
  SolveOut R = solve(P, *cfg, w0, t0, t1, ts, T, B, out, &nb);
  if (status) *status = R.status;
  if (nsteps) *nsteps = R.nsteps;
  if (naccept) *naccept = R.naccept;
  return GX_OK;
}

}  // extern "C"
