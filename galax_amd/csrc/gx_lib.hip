// gx_lib.hip -- potential handles and field evaluation kernels of
// libgalaxhip.so (declared in include/galax_amd.h).
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "gx_internal.hpp"
#include "gx_tableau.hpp"

namespace gx {

LaunchShape& launch_shape() {
  static LaunchShape s;
  return s;
}

int num_cus(int device) {
  static int cached[64] = {0};
  if (device >= 0 && device < 64 && cached[device] > 0) return cached[device];
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess ||
      n <= 0)
    n = 256;
  if (device >= 0 && device < 64) cached[device] = n;
  return n;
}

namespace {

constexpr int kFieldBlock = 256;

bool finite_all(const double* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// Gauss-Legendre nodes and weights on [0, 1] (order n): Newton's method
// on P_n from the guesses cos(pi (k - 1/4) / (n + 1/2)), P_n and P_n' by
// the three-term recurrence; weights 2 / ((1 - x^2) P_n'(x)^2), halved for
// the interval change.
void gauss_legendre01(int n, std::vector<double>& s, std::vector<double>& w) {
  s.assign(n, 0.0);
  w.assign(n, 0.0);
  auto legendre = [n](double x, double& pn, double& dpn) {
    double a = 1.0, b = x;
    for (int j = 2; j <= n; ++j) {
      const double c = ((2 * j - 1) * x * b - (j - 1) * a) / j;
      a = b;
      b = c;
    }
    pn = n == 0 ? 1.0 : b;
    const double pm = n == 1 ? 1.0 : a;
    dpn = n * (x * pn - pm) / ((x - 1.0) * (x + 1.0));
  };
  for (int k = 0; k < n; ++k) {
    double x = std::cos(M_PI * (k + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 64; ++it) {
      legendre(x, pn, dpn);
      const double dx = pn / dpn;
      x -= dx;
      if (std::fabs(dx) <= 1e-16) break;
    }
    legendre(x, pn, dpn);
    s[k] = 0.5 * (x + 1.0);
    w[k] = 1.0 / ((1.0 - x) * (1.0 + x) * dpn * dpn);
  }
}

// psi(x), the digamma function, for x > 0: the recurrence psi(x) =
// psi(x + 1) - 1/x up to x >= 20, then the asymptotic series
// ln x - 1/(2x) - sum_n B_2n / (2n x^2n) through n = 6 (truncation below
// 1e-17 there; Abramowitz & Stegun 6.3.5 and 6.3.18).
double digamma(double x) {
  double acc = 0.0;
  while (x < 20.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  const double i2 = 1.0 / (x * x);
  const double series =
      i2 * (1.0 / 12 - i2 * (1.0 / 120 - i2 * (1.0 / 252 - i2 * (1.0 / 240 - i2 * (1.0 / 132 - i2 * 691.0 / 32760)))));
  return acc + std::log(x) - 0.5 / x - series;
}

// Derived device constants for one component (see DComp in gx_potential.hpp).
// Components with a table (the triaxial NFW's quadrature nodes) append it to
// `aux` and keep its offset in p[].
int derive(const gx_component& in, DComp& out, std::vector<double>& aux) {
  std::memset(&out, 0, sizeof(out));
  out.kind = in.kind;
  out.frame = in.frame;
  out.track = in.track;
  if (!finite_all(in.p, 8) || !finite_all(in.f, 6)) return GX_EINVAL;
  const double* p = in.p;
  switch (in.kind) {
    case GX_NULL:
      break;
    case GX_KEPLER:
      out.p[0] = p[0];
      break;
    case GX_HERNQUIST:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      break;
    case GX_MIYAMOTO_NAGAI:
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = p[2] * p[2];
      break;
    case GX_NFW:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = 1.0 / p[1];
      break;
    case GX_PLUMMER:
      out.p[0] = p[0];
      out.p[1] = p[1] * p[1];
      break;
    case GX_MONARI_BAR: {
      const double alpha = p[0], R0 = p[1], v0 = p[2], Rb = p[3], phib = p[4],
                   Om = p[5];
      if (!(Rb > 0)) return GX_EINVAL;
      const double q = R0 / Rb;
      out.p[0] = alpha * (v0 * v0 / 3.0) * (q * q * q);
      out.p[1] = Rb;
      out.p[2] = 1.0 / Rb;
      out.p[3] = Rb * Rb * Rb;
      out.p[4] = 1.0 / (Rb * Rb * Rb);
      out.p[5] = 2.0 * phib;
      out.p[6] = Om;
      out.f[4] = std::cos(2.0 * phib);
      out.f[5] = std::sin(2.0 * phib);
      break;
    }
    case GX_ISOCHRONE:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = p[1] * p[1];
      break;
    case GX_JAFFE:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      break;
    case GX_LOGARITHMIC:
      out.p[0] = p[0] * p[0];
      out.p[1] = p[1] * p[1];
      break;
    case GX_LMJ09_LOG:
      if (!(p[2] > 0 && p[3] > 0 && p[4] > 0)) return GX_EINVAL;
      out.p[0] = p[0] * p[0];
      out.p[1] = p[1] * p[1];
      out.p[2] = 1.0 / (p[2] * p[2]);
      out.p[3] = 1.0 / (p[3] * p[3]);
      out.p[4] = 1.0 / (p[4] * p[4]);
      out.p[5] = std::cos(p[5]);
      out.p[6] = std::sin(p[5]);
      break;
    case GX_LONGMURALI_BAR:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0] / (2.0 * p[1]);
      out.p[1] = p[1];
      out.p[2] = p[2];
      out.p[3] = p[3] * p[3];
      out.p[4] = std::cos(p[4]);
      out.p[5] = std::sin(p[4]);
      break;
    case GX_POWERLAW_CUTOFF: {
      const double alpha = p[1], rc = p[2];
      if (!(alpha >= 0.0 && alpha < 2.0 && rc > 0)) return GX_EINVAL;
      const double a1 = 1.5 - 0.5 * alpha, a0 = 1.0 - 0.5 * alpha;
      out.p[0] = p[0];
      out.p[1] = a1;
      out.p[2] = 1.0 / (rc * rc);
      out.p[3] = std::lgamma(a1);
      out.p[4] = a0;
      out.p[5] = std::exp(std::lgamma(a0) - std::lgamma(a1)) / rc;
      out.p[6] = std::lgamma(a0);
      out.p[7] = rc;
      break;
    }
    case GX_BURKERT:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0] / (3.0 * std::log(2.0) - M_PI / 2.0);
      out.p[1] = p[1];
      out.p[2] = 1.0 / p[1];
      break;
    case GX_KUZMIN:
      out.p[0] = p[0];
      out.p[1] = p[1];
      break;
    case GX_SATOH:
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = p[2] * p[2];
      break;
    case GX_STONE:
      if (!(p[1] > 0 && p[2] > p[1])) return GX_EINVAL;
      out.p[0] = 2.0 * p[0] / (M_PI * (p[2] - p[1]));
      out.p[1] = p[1];
      out.p[2] = p[2];
      break;
    case GX_TRIAXIAL_HERNQUIST:
      if (!(p[2] > 0 && p[3] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = 1.0 / (p[2] * p[2]);
      out.p[3] = 1.0 / (p[3] * p[3]);
      break;
    case GX_TRIAXIAL_NFW: {
      const double rs = p[1], q1 = p[2], q2 = p[3];
      const int n = (int)p[4];
      if (!(rs > 0 && q1 > 0 && q2 > 0) || p[4] != (double)n || n < 1 ||
          n > GX_MAX_GL_ORDER)
        return GX_EINVAL;
      std::vector<double> sk, wk;
      gauss_legendre01(n, sk, wk);
      out.p[0] = p[0] * q1 * q2;
      out.p[1] = rs;
      out.p[2] = (double)aux.size();
      out.p[3] = n;
      for (int k = 0; k < n; ++k) {
        const double s2 = sk[k] * sk[k];
        const double A1 = 1.0 + (q1 * q1 - 1.0) * s2, A2 = 1.0 + (q2 * q2 - 1.0) * s2;
        aux.push_back(s2);
        aux.push_back(s2 / A1);
        aux.push_back(s2 / A2);
        aux.push_back(wk[k] / std::sqrt(A1 * A2));
      }
      break;
    }
    case GX_LEESUTO_NFW: {
      const double rs = p[1], a1 = p[2], a2 = p[3], a3 = p[4];
      if (!(rs > 0 && a3 > 0 && a1 >= a2 && a2 >= a3)) return GX_EINVAL;
      const double eb2 = 1.0 - (a2 / a1) * (a2 / a1), ec2 = 1.0 - (a3 / a1) * (a3 / a1);
      const double ln2 = std::log(2.0);
      out.p[0] = (p[0] / rs) / (ln2 - 0.5 + (ln2 - 0.75) * (eb2 + ec2));
      out.p[1] = 1.0 / rs;
      out.p[2] = eb2;
      out.p[3] = ec2;
      out.p[4] = 0.5 * (eb2 + ec2);
      break;
    }
    case GX_VOGELSBERGER_NFW: {
      const double rs = p[1], q1sq = p[2] * p[2];
      if (!(rs > 0 && q1sq > 0 && q1sq < 3.0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = rs;
      out.p[2] = 1.0 / q1sq;
      out.p[3] = 1.0 / (3.0 - q1sq);
      out.p[4] = p[3] * rs;
      break;
    }
    case GX_HARMONIC_OSCILLATOR:
      out.p[0] = p[0] * p[0];
      break;
    case GX_HENON_HEILES:
      if (!(p[1] != 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = 1.0 / (p[1] * p[1]);
      break;
    case GX_HARDCUTOFF_NFW: {
      const double rs = p[1], rt = p[2];
      if (!(rs > 0 && rt > 0)) return GX_EINVAL;
      const double xt = rt / rs;
      out.p[0] = p[0];
      out.p[1] = rs;
      out.p[2] = 1.0 / rs;
      out.p[3] = p[0] * (std::log1p(xt) - xt / (1.0 + xt));  // G M_NFW(<r_t)
      out.p[4] = rt;
      break;
    }
    case GX_GNFW: {
      const double rs = p[1], gamma = p[2];
      if (!(rs > 0 && gamma >= 0.0 && gamma < 2.0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = rs;
      out.p[2] = 1.0 / rs;
      out.p[3] = 3.0 - gamma;
      out.p[4] = digamma(3.0 - gamma);
      break;
    }
    case GX_GAUSSIAN:
      if (!(p[1] > 0)) return GX_EINVAL;
      out.p[0] = p[0];
      out.p[1] = p[1];
      out.p[2] = 1.0 / p[1];
      break;
    case GX_TRIAXIAL_GAUSSIAN: {
      // the triaxial NFW's node table [s^2, s^2/A1, s^2/A2, w/sqrt(A1 A2)]
      const double rs = p[1], q1 = p[2], q2 = p[3];
      const int n = (int)p[4];
      if (!(rs > 0 && q1 > 0 && q2 > 0) || p[4] != (double)n || n < 1 ||
          n > GX_MAX_GL_ORDER)
        return GX_EINVAL;
      std::vector<double> sk, wk;
      gauss_legendre01(n, sk, wk);
      out.p[0] = 2.0 * p[0] / (std::sqrt(2.0 * M_PI) * rs * rs * rs);
      out.p[1] = 1.0 / (rs * rs);
      out.p[2] = (double)aux.size();
      out.p[3] = n;
      out.p[4] = rs * rs;
      for (int k = 0; k < n; ++k) {
        const double s2 = sk[k] * sk[k];
        const double A1 = 1.0 + (q1 * q1 - 1.0) * s2, A2 = 1.0 + (q2 * q2 - 1.0) * s2;
        aux.push_back(s2);
        aux.push_back(s2 / A1);
        aux.push_back(s2 / A2);
        aux.push_back(wk[k] / std::sqrt(A1 * A2));
      }
      break;
    }
    default:
      return GX_EINVAL;
  }
  // raw table parameters for the adjoint (spare slots of the basic kinds)
  if (in.kind >= GX_KEPLER && in.kind <= GX_PLUMMER) {
    out.p[5] = p[0];
    out.p[6] = p[1];
    out.p[7] = p[2];
  }
  switch (in.frame) {
    case GX_FRAME_NONE:
      break;
    case GX_FRAME_OFFSET:
    case GX_FRAME_ROTATING:
      for (int k = 0; k < 4; ++k) out.f[k] = in.f[k];
      break;
    case GX_FRAME_TRACK:
      break;
    default:
      return GX_EINVAL;
  }
  return GX_OK;
}

// Rotation rate that lets a component join a co-rotating group, or NaN.
double group_rate(const gx_component& c) {
  if (c.frame == GX_FRAME_ROTATING) return c.f[3];
  if (c.kind == GX_MONARI_BAR && c.frame == GX_FRAME_NONE) return c.p[5];
  return NAN;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

__global__ __launch_bounds__(kFieldBlock) void k_accel(
    DPot P, const double* __restrict__ xyz, const double* __restrict__ t,
    int t_stride, long long n, double* __restrict__ out, double sign) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const double ti = t[t_stride ? i : 0];
    double ax, ay, az;
    accel<true>(P, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ti, ax, ay,
                 az);
    out[3 * i] = sign * ax;
    out[3 * i + 1] = sign * ay;
    out[3 * i + 2] = sign * az;
  }
}

__global__ __launch_bounds__(kFieldBlock) void k_potential(
    DPot P, const double* __restrict__ xyz, const double* __restrict__ t,
    int t_stride, long long n, double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const double ti = t[t_stride ? i : 0];
    out[i] = potential(P, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ti);
  }
}

// The short FP64 primitives of gx_math.hpp at given arguments (accuracy
// checks), GX_MATH_CHECK_OUTPUTS values per argument in the order the
// header lists them.
__global__ void k_math_check(const double* __restrict__ x, long long n,
                             double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  double* o = out + GX_MATH_CHECK_OUTPUTS * i;
  o[0] = rcp(v);
  o[1] = rsqrt_fast(v);
  o[2] = log1p_fast(v);
  o[3] = div(1.0, v);
  o[4] = sqrt_fast(v);
  o[5] = div6(v);
  o[6] = root_n<9>(v);
  o[7] = root_n<6>(v);
  o[8] = rcp(rsqrt_fast(rsqrt_fast(rsqrt_fast(v))));
}

__global__ __launch_bounds__(kFieldBlock) void k_hessian(
    DPot P, const double* __restrict__ xyz, const double* __restrict__ t,
    int t_stride, long long n, double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const double ti = t[t_stride ? i : 0];
    double H[9];
    hessian(P, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ti, H);
    for (int k = 0; k < 9; ++k) out[9 * i + k] = H[k];
  }
}

int field_grid(const gx_pot* pot, long long n) {
  long long blocks = (n + kFieldBlock - 1) / kFieldBlock;
  long long cap = (long long)num_cus(pot->device) * 16;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

int field_args_ok(const gx_pot* pot, const void* xyz, const void* t,
                  int t_stride, long long n, const void* out) {
  if (!pot || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!xyz || !t || !out || (t_stride != 0 && t_stride != 1)) return GX_EINVAL;
  return 1;
}

// LDS that a launch may give a staged dense-solution track (pieces and
// cells); the integrator's own shared state takes about 22 KB of the
// workgroup's 160 KB.
constexpr long long kLdsTrackBytes = 128 * 1024;

// Pieces, cells and boundaries of a dense-solution track (the layout that
// DTrack documents) from n accepted-step records of method M, in blob;
// fills d's interval, end positions and cell geometry (not its pointers).
// The records may run forward or backward in time; every step must have
// the sign of the first, and each must start where the previous one ended
// (to rounding).
template <class M>
int build_pieces(const gx_track& tr, std::vector<double>& blob, DTrack& d) {
  constexpr int SF = M::SF;
  constexpr int D = sizeof(M::dn.wq[0]) / sizeof(double);
  static_assert(D + 1 <= 7, "pieces hold degree-6 polynomials");
  const long long n = tr.n;
  const double* R = tr.rec;
  for (long long k = 0; k < n * GX_DENSE_RECORD; ++k)
    if (!std::isfinite(R[k])) return GX_EINVAL;
  const bool fwd = R[1] > 0.0;
  std::vector<double> brk(n + 1), right(n);
  std::vector<double> pieces(kPieceDoubles * n, 0.0);
  for (long long j = 0; j < n; ++j) {
    const double* r = R + GX_DENSE_RECORD * (fwd ? j : n - 1 - j);
    const double ts = r[0], h = r[1];
    if (!(fwd ? h > 0.0 : h < 0.0)) return GX_EINVAL;
    double* pc = pieces.data() + kPieceDoubles * j;
    pc[0] = ts;
    pc[1] = 1.0 / h;
    // q(th) = q + th h v + h^2 th sum_l Wq_l(th) g_l, Wq in Horner order
    // (th^(D-1) first): the power of th of wq[l][e] is D - e
    for (int m = 0; m < 3; ++m) {
      double C[7] = {r[2 + m], h * r[5 + m], 0, 0, 0, 0, 0};
      for (int e = 0; e < D; ++e) {
        double acc = 0.0;
        for (int l = 0; l < SF; ++l) acc += M::dn.wq[l][e] * r[8 + 3 * l + m];
        C[D - e] += h * h * acc;
      }
      for (int e = 0; e < 7; ++e) pc[2 + 7 * m + e] = C[e];
    }
    brk[j] = fwd ? ts : ts + h;
    right[j] = fwd ? ts + h : ts;
  }
  brk[n] = right[n - 1];
  for (long long j = 0; j + 1 < n; ++j) {
    // the end of piece j against the start of piece j + 1
    const double tol = 64.0 * std::numeric_limits<double>::epsilon() *
                       std::fmax(1.0, std::fabs(right[j]));
    if (!(brk[j + 1] > brk[j]) || std::fabs(right[j] - brk[j + 1]) > tol)
      return GX_EINVAL;
  }
  if (!(brk[n] > brk[n - 1])) return GX_EINVAL;
  auto eval = [&](long long j, double t, double* q) {
    const double* pc = pieces.data() + kPieceDoubles * j;
    const double th = (t - pc[0]) * pc[1];
    for (int m = 0; m < 3; ++m) {
      const double* a = pc + 2 + 7 * m;
      double v = a[6];
      for (int e = 5; e >= 0; --e) v = std::fma(v, th, a[e]);
      q[m] = v;
    }
  };
  d.kind = kTrackPieces;
  d.n = n;
  d.t0 = brk[0];
  d.t1 = brk[n];
  eval(0, d.t0, d.q0);
  eval(n - 1, d.t1, d.q1);
  // cells: the coarsest table (2n, 4n, ... cells) in which every cell's
  // times, widened by a millionth of a cell either side for rounding,
  // cross at most two piece starts; none when LDS cannot hold one
  const double inf = std::numeric_limits<double>::infinity();
  const double span = d.t1 - d.t0;
  std::vector<double> cells;
  int ncell = 0;
  for (long long nc = 2 * n; nc <= 64 * n; nc *= 2) {
    if (8 * (kPieceDoubles * n + kCellDoubles * nc) > kLdsTrackBytes || nc > INT32_MAX) break;
    cells.assign(kCellDoubles * nc, 0.0);
    const double w = span / (double)nc, slack = 1e-6 * w;
    bool ok = true;
    long long k = 0;
    for (long long c = 0; c < nc && ok; ++c) {
      const double a = d.t0 + (double)c * w - slack, b = d.t0 + (double)(c + 1) * w + slack;
      while (k + 1 < n && brk[k + 1] <= a) ++k;
      ok = k + 3 >= n || brk[k + 3] > b;
      double* e = cells.data() + kCellDoubles * c;
      e[0] = (double)k;
      e[1] = k + 1 < n ? brk[k + 1] : inf;
      e[2] = k + 2 < n ? brk[k + 2] : inf;
    }
    if (ok) {
      ncell = (int)nc;
      break;
    }
  }
  if (ncell == 0) cells.clear();
  d.ncell = ncell;
  d.inv_h = ncell > 0 ? (double)ncell / span : 0.0;
  d.lds_doubles = ncell > 0 ? kPieceDoubles * n + kCellDoubles * (long long)ncell : 0;
  blob = std::move(pieces);
  blob.insert(blob.end(), cells.begin(), cells.end());
  blob.insert(blob.end(), brk.begin(), brk.end());
  return GX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gx_potential_create(const gx_component* comps, int32_t ncomp,
                        const gx_track* tracks, int32_t ntracks,
                        gx_pot** out) {
  if (!out || ncomp < 0 || ncomp > GX_MAX_COMPONENTS || ntracks < 0 ||
      ntracks > GX_MAX_TRACKS || (ncomp > 0 && !comps) ||
      (ntracks > 0 && !tracks))
    return GX_EINVAL;
  *out = nullptr;
  std::vector<DComp> dc(ncomp > 0 ? ncomp : 1);
  std::vector<double> aux;
  bool has_track = false, has_bar = false;
  for (int i = 0; i < ncomp; ++i) {
    if (comps[i].reserved != 0) return GX_EINVAL;
    if (derive(comps[i], dc[i], aux) != GX_OK) return GX_EINVAL;
    if (comps[i].frame == GX_FRAME_TRACK) {
      if (comps[i].track < 0 || comps[i].track >= ntracks) return GX_EINVAL;
      has_track = true;
    }
    if (comps[i].kind == GX_MONARI_BAR) has_bar = true;
  }
  // co-rotating groups: runs of consecutive components with the same rate
  for (int i = 0; i < ncomp;) {
    const double w = group_rate(comps[i]);
    if (std::isnan(w)) {
      ++i;
      continue;
    }
    int j = i + 1;
    while (j < ncomp && group_rate(comps[j]) == w) ++j;
    for (int k = i; k < j; ++k) {
      dc[k].flags = kInGroup;
      dc[k].f[3] = w;
    }
    dc[i].flags |= kGroupStart;
    dc[j - 1].flags |= kGroupEnd;
    i = j;
  }
  // tracks: validate knots; build dense-solution tracks
  size_t knot_bytes = 0;
  std::vector<std::vector<double>> blobs(ntracks > 0 ? ntracks : 1);
  std::vector<DTrack> built(ntracks > 0 ? ntracks : 1);
  for (int k = 0; k < ntracks; ++k) {
    const gx_track& tr = tracks[k];
    if (tr.reserved != 0) return GX_EINVAL;
    if (tr.rec) {
      if (tr.n < 1 || tr.n > INT32_MAX) return GX_EINVAL;
      built[k] = DTrack{};
      const int rc = tr.method == GX_DOPRI8   ? build_pieces<Dopri8>(tr, blobs[k], built[k])
                     : tr.method == GX_DOPRI5 ? build_pieces<Dopri5>(tr, blobs[k], built[k])
                                              : GX_EINVAL;
      if (rc != GX_OK) return rc;
      knot_bytes += align256(sizeof(double) * blobs[k].size());
      continue;
    }
    if (tr.n < 1 || tr.n > INT32_MAX || !tr.t || !tr.q || !tr.v) return GX_EINVAL;
    for (int64_t i = 1; i < tr.n; ++i)
      if (!(tr.t[i] > tr.t[i - 1])) return GX_EINVAL;
    for (int64_t i = 0; i < tr.n; ++i)
      if (!std::isfinite(tr.t[i]) || !std::isfinite(tr.q[3 * i]) ||
          !std::isfinite(tr.q[3 * i + 1]) || !std::isfinite(tr.q[3 * i + 2]) ||
          !std::isfinite(tr.v[3 * i]) || !std::isfinite(tr.v[3 * i + 1]) ||
          !std::isfinite(tr.v[3 * i + 2]))
        return GX_EINVAL;
    knot_bytes += align256(sizeof(double) * 16 * (tr.n > 1 ? tr.n - 1 : 1));
  }
  const size_t comp_bytes = align256(sizeof(DComp) * dc.size());
  const size_t track_bytes = align256(sizeof(DTrack) * (ntracks > 0 ? ntracks : 1));
  const size_t aux_bytes = align256(sizeof(double) * (aux.empty() ? 1 : aux.size()));
  const size_t total = comp_bytes + track_bytes + knot_bytes + aux_bytes;

  gx_pot* h = new gx_pot();
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return GX_ELAUNCH;
  }
  if (hipMalloc(&h->block, total) != hipSuccess) {
    delete h;
    return GX_ENOMEM;
  }
  char* base = static_cast<char*>(h->block);
  std::vector<char> host(total, 0);
  std::memcpy(host.data(), dc.data(), sizeof(DComp) * dc.size());
  std::vector<DTrack> dt(ntracks > 0 ? ntracks : 1);
  size_t off = comp_bytes + track_bytes;
  for (int k = 0; k < ntracks; ++k) {
    const gx_track& tr = tracks[k];
    if (tr.rec) {
      DTrack d = built[k];
      std::memcpy(host.data() + off, blobs[k].data(), sizeof(double) * blobs[k].size());
      d.rec = reinterpret_cast<const double*>(base + off);
      d.cells = d.rec + kPieceDoubles * d.n;
      d.brk = d.cells + kCellDoubles * (long long)d.ncell;
      off += align256(sizeof(double) * blobs[k].size());
      dt[k] = d;
      continue;
    }
    DTrack d{};
    d.kind = kTrackKnots;
    d.n = tr.n;
    d.t0 = tr.t[0];
    d.t1 = tr.t[tr.n - 1];
    d.inv_h = tr.n > 1 ? (double)(tr.n - 1) / (tr.t[tr.n - 1] - tr.t[0]) : 0.0;
    for (int m = 0; m < 3; ++m) {
      d.q0[m] = tr.q[m];
      d.q1[m] = tr.q[3 * (tr.n - 1) + m];
    }
    // cubic Hermite interpolant of (q, v) at the knots in power form:
    // c0 = q_k, c1 = h v_k, c2 = 3 (q_{k+1} - q_k) - h (2 v_k + v_{k+1}),
    // c3 = 2 (q_k - q_{k+1}) + h (v_k + v_{k+1})
    double* rec = reinterpret_cast<double*>(host.data() + off);
    for (int64_t i = 0; i + 1 < tr.n; ++i) {
      double* r = rec + 16 * i;
      const double hk = tr.t[i + 1] - tr.t[i];
      r[0] = tr.t[i];
      r[1] = tr.t[i + 1];
      r[2] = 1.0 / hk;
      r[3] = 0.0;
      for (int m = 0; m < 3; ++m) {
        const double qa = tr.q[3 * i + m], qb = tr.q[3 * (i + 1) + m];
        const double va = hk * tr.v[3 * i + m], vb = hk * tr.v[3 * (i + 1) + m];
        r[4 + 4 * m] = qa;
        r[5 + 4 * m] = va;
        r[6 + 4 * m] = 3.0 * (qb - qa) - 2.0 * va - vb;
        r[7 + 4 * m] = 2.0 * (qa - qb) + va + vb;
      }
    }
    d.rec = reinterpret_cast<const double*>(base + off);
    off += align256(sizeof(double) * 16 * (tr.n > 1 ? tr.n - 1 : 1));
    dt[k] = d;
  }
  std::memcpy(host.data() + comp_bytes, dt.data(), sizeof(DTrack) * dt.size());
  if (!aux.empty())
    std::memcpy(host.data() + off, aux.data(), sizeof(double) * aux.size());
  if (hipMemcpy(h->block, host.data(), total, hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(h->block);
    delete h;
    return GX_ELAUNCH;
  }
  h->dev.comps = reinterpret_cast<const DComp*>(base);
  h->dev.tracks = reinterpret_cast<const DTrack*>(base + comp_bytes);
  h->dev.aux = reinterpret_cast<const double*>(base + off);
  h->dev.n = ncomp;
  h->dev.ntracks = ntracks;
  h->ncomp = ncomp;
  h->has_track = has_track;
  h->has_bar = has_bar;
  h->basic = true;
  for (int i = 0; i < ncomp; ++i)
    if (comps[i].frame != GX_FRAME_NONE || comps[i].kind == GX_MONARI_BAR)
      h->basic = false;
  h->kind0 = ncomp > 0 ? comps[0].kind : GX_NULL;
  h->sig = kSigNone;
  if (h->basic && ncomp == 4 && comps[0].kind == GX_MIYAMOTO_NAGAI &&
      comps[1].kind == GX_NFW && comps[2].kind == GX_HERNQUIST &&
      comps[3].kind == GX_HERNQUIST)
    h->sig = kSigMilkyWay;
  bool mw_origin = ncomp == 5;
  for (int i = 0; mw_origin && i < 4; ++i) mw_origin = comps[i].frame == GX_FRAME_NONE;
  if (mw_origin && comps[0].kind == GX_MIYAMOTO_NAGAI && comps[1].kind == GX_NFW &&
      comps[2].kind == GX_HERNQUIST && comps[3].kind == GX_HERNQUIST &&
      comps[4].kind == GX_PLUMMER && comps[4].frame == GX_FRAME_TRACK) {
    // a knot track, or a dense-solution track that LDS holds with its
    // cells (otherwise the table evaluator, which looks pieces up in
    // device memory)
    const DTrack& tk = dt[comps[4].track];
    h->sig = tk.kind == kTrackKnots ? kSigMilkyWayTracked
             : tk.ncell > 0        ? kSigMilkyWayPieces
                                   : kSigNone;
    h->dev.trk = tk;
    h->dev.kp[0] = dc[4].p[0];
    h->dev.kp[1] = dc[4].p[1];
  }
  bool rot = mw_origin || ncomp > 5;
  for (int i = 0; rot && i < 4; ++i) rot = comps[i].frame == GX_FRAME_NONE;
  rot = rot && ncomp >= 5 && comps[0].kind == GX_MIYAMOTO_NAGAI && comps[1].kind == GX_NFW &&
        comps[2].kind == GX_HERNQUIST && comps[3].kind == GX_HERNQUIST &&
        (dc[4].flags & kGroupStart) && (dc[ncomp - 1].flags & kGroupEnd);
  for (int i = 4; rot && i < ncomp; ++i) {
    const bool bar = comps[i].kind == GX_MONARI_BAR && comps[i].frame == GX_FRAME_NONE;
    const bool sub = comps[i].kind == GX_PLUMMER && comps[i].frame == GX_FRAME_ROTATING;
    rot = (bar || sub) && (dc[i].flags & kInGroup) && (i == 4 || !(dc[i].flags & kGroupStart));
  }
  if (rot) h->sig = kSigMilkyWayRotating;
  *out = h;
  return GX_OK;
}

int gx_potential_destroy(gx_pot* pot) {
  if (!pot) return GX_EINVAL;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(pot->device);
  (void)hipFree(pot->block);
  (void)hipSetDevice(prev);
  delete pot;
  return GX_OK;
}

static int launch_accel(const gx_pot* pot, const double* xyz, const double* t,
                        int32_t t_stride, int64_t n, double* out, void* stream,
                        double sign) {
  int ok = field_args_ok(pot, xyz, t, t_stride, n, out);
  if (ok != 1) return ok;
  hipLaunchKernelGGL(k_accel, dim3(field_grid(pot, n)), dim3(kFieldBlock), 0,
                     (hipStream_t)stream, pot->dev, xyz, t, t_stride,
                     (long long)n, out, sign);
  return check(hipGetLastError());
}

int gx_acceleration(const gx_pot* pot, const double* xyz, const double* t,
                    int32_t t_stride, int64_t n, double* acc, void* stream) {
  return launch_accel(pot, xyz, t, t_stride, n, acc, stream, 1.0);
}

int gx_gradient(const gx_pot* pot, const double* xyz, const double* t,
                int32_t t_stride, int64_t n, double* grad, void* stream) {
  return launch_accel(pot, xyz, t, t_stride, n, grad, stream, -1.0);
}

int gx_potential_eval(const gx_pot* pot, const double* xyz, const double* t,
                      int32_t t_stride, int64_t n, double* phi, void* stream) {
  int ok = field_args_ok(pot, xyz, t, t_stride, n, phi);
  if (ok != 1) return ok;
  hipLaunchKernelGGL(k_potential, dim3(field_grid(pot, n)), dim3(kFieldBlock),
                     0, (hipStream_t)stream, pot->dev, xyz, t, t_stride,
                     (long long)n, phi);
  return check(hipGetLastError());
}

int gx_hessian(const gx_pot* pot, const double* xyz, const double* t,
               int32_t t_stride, int64_t n, double* hess, void* stream) {
  int ok = field_args_ok(pot, xyz, t, t_stride, n, hess);
  if (ok != 1) return ok;
  hipLaunchKernelGGL(k_hessian, dim3(field_grid(pot, n)), dim3(kFieldBlock), 0,
                     (hipStream_t)stream, pot->dev, xyz, t, t_stride,
                     (long long)n, hess);
  return check(hipGetLastError());
}

int gx_set_launch_shape(int32_t block, int32_t blocks_per_cu,
                        int32_t* prev_block, int32_t* prev_blocks_per_cu) {
  LaunchShape& s = launch_shape();
  if (prev_block) *prev_block = s.block;
  if (prev_blocks_per_cu) *prev_blocks_per_cu = s.blocks_per_cu;
  if (block != 0 && block != s.block) return GX_EINVAL;  // fixed at 256
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return GX_EINVAL;
  if (blocks_per_cu > 0) s.blocks_per_cu = blocks_per_cu;
  return GX_OK;
}

int gx_set_launch_fraction(int32_t parts, int32_t* prev_parts) {
  LaunchShape& s = launch_shape();
  if (prev_parts) *prev_parts = s.parts;
  if (parts < 0 || parts > 64) return GX_EINVAL;
  if (parts > 0) s.parts = parts;
  return GX_OK;
}

int gx_set_queue_policy(int64_t head_lanes, int32_t chunk,
                        int64_t* prev_head_lanes, int32_t* prev_chunk) {
  LaunchShape& s = launch_shape();
  if (prev_head_lanes) *prev_head_lanes = s.head_lanes;
  if (prev_chunk) *prev_chunk = s.chunk;
  if (head_lanes > ((int64_t)1 << 30) || chunk < 0 || chunk > 64)
    return GX_EINVAL;
  if (head_lanes >= 0) s.head_lanes = head_lanes;
  if (chunk > 0) s.chunk = chunk;
  return GX_OK;
}

int gx_math_check(const double* x, int64_t n, double* out, void* stream) {
  if (!x || !out || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if ((n + 255) / 256 > 0x7fffffffLL) return GX_EINVAL;
  hipLaunchKernelGGL(k_math_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, x, (long long)n, out);
  return check(hipGetLastError());
}

int gx_device_info(int32_t* num_cu, int32_t* warp_size) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GX_ELAUNCH;
  if (num_cu) *num_cu = num_cus(dev);
  if (warp_size) {
    int w = 64;
    (void)hipDeviceGetAttribute(&w, hipDeviceAttributeWarpSize, dev);
    *warp_size = w;
  }
  return GX_OK;
}

}  // extern "C"
