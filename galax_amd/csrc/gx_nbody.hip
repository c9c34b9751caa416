// gx_nbody.hip -- direct-summation N-body field and its integrator.
//
// galax's NBodyField (dynamics/_src/orbit/field_nbody.py:94-131):
//   a_j = sum_{i != j} G m_i (x_i - x_j) / (|x_i - x_j|^2 + eps^2)^(3/2)
//         - grad Phi_ext(x_j, t),
// integrated as ONE system by OrbitSolver (shared steps; the PID error norm
// is the RMS over all 6N state components, orbit/solver.py:121-141).
//
// One workgroup of 512 threads integrates the whole system (N <= 4096, the
// range in which the reference's (N, N, 3) pair arrays are practical): the
// stage positions of all bodies sit in LDS, every thread sums the pull of
// all bodies on its own bodies (all lanes read the same LDS word: a
// broadcast), stage accelerations go to a workspace in HBM, and the error
// norm is a block reduction.  Control flow (step accept/reject, saves) is
// uniform across the workgroup, so no host round trip happens per step.
#include <hip/hip_runtime.h>

#include <cmath>

#include "gx_internal.hpp"
#include "gx_potential.hpp"
#include "gx_rk.hpp"
#include "gx_tableau.hpp"

namespace gx {
namespace {

constexpr int kNbThreads = 512;
constexpr int kNbMax = 4096;

struct NArgs {
  double dt0, rtol, atol, safety, factormin, factormax, dtmin, dtmax;
  double c1, c2, c3;
  long long max_steps;
  int force_dtmin;
  int T;
  double eps2;
  double t0, t1;
  const double* gm;  // [n] G m
  const double* w0;  // [n][6]
  const double* ts;  // [T]
  long long n;
  double* out;       // [n][T][6]
  int* status;
  long long* nsteps;
  long long* naccept;
  double* q;         // workspace: state [n][3] (positions, solver-time velocities)
  double* p;
  double* q1;        // end of the current step [n][3]
  double* p1;
  double* G;         // stage accelerations [SF][n][3] (slot SF-1: the end stage)
};

struct NbShared {
  double x[kNbMax], y[kNbMax], z[kNbMax], gm[kNbMax];
  double red[kNbThreads / 64];
};

__device__ __forceinline__ double block_sum(NbShared& S, double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) S.red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < kNbThreads / 64; ++i) s += S.red[i];
  return s;
}

// Acceleration of body j at (x, y, z) from the positions in LDS and the
// external field at physical time t.
__device__ __forceinline__ void nb_accel(const DPot& P, const NbShared& S, long long n,
                                         double eps2, long long j, double x, double y,
                                         double z, double t, double& ax, double& ay,
                                         double& az) {
  accel<true>(P, x, y, z, t, ax, ay, az);
  for (long long k = 0; k < n; ++k) {
    const double dx = S.x[k] - x, dy = S.y[k] - y, dz = S.z[k] - z;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    const double ir = rsqrt_fast(r2);
    const double w = (k == j) ? 0.0 : S.gm[k] * ir * ir * ir;
    ax = fma(w, dx, ax);
    ay = fma(w, dy, ay);
    az = fma(w, dz, az);
  }
}

// One method evaluation for the whole system from (q, p) at solver time tau
// with step h: writes q1, p1 and (ERR) returns the squared scaled error sum
// of this thread's bodies.
template <class M, bool ERR>
__device__ double nb_eval(const DPot& P, NbShared& S, const NArgs& A, double dir,
                          double tau, double h) {
  constexpr auto& K = M::k;
  const long long n = A.n;
  const int tid = threadIdx.x;
  for (int i = 0; i < M::S; ++i) {
    const double ti = dir * fma(K.c[i], h, tau);
    // stage positions of own bodies into LDS
    for (long long j = tid; j < n; j += kNbThreads) {
      double s[3] = {0.0, 0.0, 0.0};
      for (int l = 0; l < i; ++l) {
        const double a = K.aa[i][l];
        if (a != 0.0)
          for (int m = 0; m < 3; ++m) s[m] = fma(a, A.G[((long long)l * n + j) * 3 + m], s[m]);
      }
      const double rs = K.rs[i];
      S.x[j] = fma(h, fma(rs, A.p[3 * j], h * s[0]), A.q[3 * j]);
      S.y[j] = fma(h, fma(rs, A.p[3 * j + 1], h * s[1]), A.q[3 * j + 1]);
      S.z[j] = fma(h, fma(rs, A.p[3 * j + 2], h * s[2]), A.q[3 * j + 2]);
    }
    __syncthreads();
    for (long long j = tid; j < n; j += kNbThreads) {
      double ax, ay, az;
      nb_accel(P, S, n, A.eps2, j, S.x[j], S.y[j], S.z[j], ti, ax, ay, az);
      double* g = A.G + ((long long)i * n + j) * 3;
      g[0] = ax;
      g[1] = ay;
      g[2] = az;
    }
    __syncthreads();
  }
  double sq = 0.0;
  for (long long j = tid; j < n; j += kNbThreads) {
    double b[3] = {0.0, 0.0, 0.0}, ba[3] = {0.0, 0.0, 0.0};
    double e[3] = {0.0, 0.0, 0.0}, ea[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < M::S; ++l) {
      for (int m = 0; m < 3; ++m) {
        const double gl = A.G[((long long)l * n + j) * 3 + m];
        if (K.b[l] != 0.0) b[m] = fma(K.b[l], gl, b[m]);
        if (K.ba[l] != 0.0) ba[m] = fma(K.ba[l], gl, ba[m]);
        if (ERR && K.e[l] != 0.0) e[m] = fma(K.e[l], gl, e[m]);
        if (ERR && K.ea[l] != 0.0) ea[m] = fma(K.ea[l], gl, ea[m]);
      }
    }
    for (int m = 0; m < 3; ++m) {
      const double q0 = A.q[3 * j + m], p0 = A.p[3 * j + m];
      const double q1 = fma(h, fma(K.sb, p0, h * ba[m]), q0);
      const double p1 = fma(h, b[m], p0);
      A.q1[3 * j + m] = q1;
      A.p1[3 * j + m] = p1;
      if (ERR) {
        const double eq = h * fma(K.se, p0, h * ea[m]);
        const double ep = h * e[m];
        const double uq = eq / (A.atol + fmax(fabs(q0), fabs(q1)) * A.rtol);
        const double up = ep / (A.atol + fmax(fabs(p0), fabs(p1)) * A.rtol);
        sq += uq * uq + up * up;
      }
    }
  }
  return sq;
}

__device__ void nb_write(const NArgs& A, int s, double dir, const double* q, const double* p) {
  for (long long j = threadIdx.x; j < A.n; j += kNbThreads) {
    double* o = A.out + (j * A.T + s) * 6;
    for (int m = 0; m < 3; ++m) {
      o[m] = q[3 * j + m];
      o[3 + m] = dir * p[3 * j + m];
    }
  }
}

template <class M, bool PID>
__global__ void __launch_bounds__(kNbThreads) k_nbody(DPot P, NArgs A) {
  __shared__ NbShared S;
  const long long n = A.n;
  const int tid = threadIdx.x;
  const double dir = (A.t0 < A.t1) ? 1.0 : -1.0;
  for (long long j = tid; j < n; j += kNbThreads) {
    S.gm[j] = A.gm[j];
    for (int m = 0; m < 3; ++m) {
      A.q[3 * j + m] = A.w0[6 * j + m];
      A.p[3 * j + m] = dir * A.w0[6 * j + 3 + m];
    }
  }
  __syncthreads();
  double tau = dir * A.t0;
  const double tau1 = dir * A.t1;
  int si = 0;
  while (si < A.T && dir * A.ts[si] <= tau) nb_write(A, si++, dir, A.q, A.p);
  int status = GX_SUCCESS;
  long long nst = 0, nacc = 0;
  if (A.t0 != A.t1) {
    // first step: dt0; dtmin if the controller sets one (as the orbit
    // integrators); else diffrax's initial-step heuristic over the system
    double h0 = A.dt0 > 0.0 ? A.dt0 : (PID && A.dtmin > 0.0 ? A.dtmin : 0.0);
    if (!(h0 > 0.0)) {
      double s0 = 0.0, s1 = 0.0;
      for (long long j = tid; j < n; j += kNbThreads) {
        S.x[j] = A.q[3 * j];
        S.y[j] = A.q[3 * j + 1];
        S.z[j] = A.q[3 * j + 2];
      }
      __syncthreads();
      for (long long j = tid; j < n; j += kNbThreads) {
        double a[3];
        nb_accel(P, S, n, A.eps2, j, S.x[j], S.y[j], S.z[j], dir * tau, a[0], a[1], a[2]);
        for (int m = 0; m < 3; ++m) {
          const double sq = A.atol + fabs(A.q[3 * j + m]) * A.rtol;
          const double sp = A.atol + fabs(A.p[3 * j + m]) * A.rtol;
          s0 += (A.q[3 * j + m] / sq) * (A.q[3 * j + m] / sq) + (A.p[3 * j + m] / sp) * (A.p[3 * j + m] / sp);
          s1 += (A.p[3 * j + m] / sq) * (A.p[3 * j + m] / sq) + (a[m] / sp) * (a[m] / sp);
          A.G[3 * j + m] = a[m];  // f0's acceleration part, kept for d2
        }
      }
      const double nn = 6.0 * (double)n;
      const double d0 = sqrt(block_sum(S, s0) / nn);
      const double d1 = sqrt(block_sum(S, s1) / nn);
      const bool cond = (d0 < 1e-5) || (d1 < 1e-5);
      const double hh = cond ? 1e-6 : 0.01 * (d0 / (cond ? 1.0 : d1));
      __syncthreads();
      for (long long j = tid; j < n; j += kNbThreads) {
        S.x[j] = A.q[3 * j] + A.p[3 * j] * hh;
        S.y[j] = A.q[3 * j + 1] + A.p[3 * j + 1] * hh;
        S.z[j] = A.q[3 * j + 2] + A.p[3 * j + 2] * hh;
      }
      __syncthreads();
      double s2 = 0.0;
      for (long long j = tid; j < n; j += kNbThreads) {
        double a[3];
        nb_accel(P, S, n, A.eps2, j, S.x[j], S.y[j], S.z[j], dir * (tau + hh), a[0], a[1], a[2]);
        for (int m = 0; m < 3; ++m) {
          const double sq = A.atol + fabs(A.q[3 * j + m]) * A.rtol;
          const double sp = A.atol + fabs(A.p[3 * j + m]) * A.rtol;
          const double p1 = A.p[3 * j + m] + dir * A.G[3 * j + m] * dir * hh;
          const double u = (p1 - A.p[3 * j + m]) / sq;
          const double v = dir * (a[m] - A.G[3 * j + m]) / sp;
          s2 += u * u + v * v;
        }
      }
      const double d2 = sqrt(block_sum(S, s2) / nn) / hh;
      const double maxd = fmax(d1, d2);
      const double h1 = (maxd <= 1e-15) ? fmax(1e-6, hh * 1e-3)
                                        : pow(0.01 / maxd, 1.0 / ((double)M::order + 1.0));
      h0 = fmin(100.0 * hh, h1);
    }
    bool at_dtmin = false;
    if (PID) {
      h0 = fmin(h0, A.dtmax);
      if (A.dtmin > 0.0) {
        at_dtmin = h0 <= A.dtmin;
        h0 = fmax(h0, A.dtmin);
      }
    }
    double tnext = fmin(tau + h0, tau1);
    double prev_inv = 1.0, prev_prev_inv = 1.0;
    while (tau < tau1) {
      if (nst >= A.max_steps) {
        status = GX_MAX_STEPS_REACHED;
        break;
      }
      const double h = tnext - tau;
      if (!(h >= 0.0)) {
        status = GX_NONFINITE;
        break;
      }
      const double sq = block_sum(S, nb_eval<M, PID>(P, S, A, dir, tau, h));
      ++nst;
      // a non-finite state anywhere stops the system
      double bad = 0.0;
      for (long long j = tid; j < n; j += kNbThreads)
        for (int m = 0; m < 3; ++m)
          if (!isfinite(A.q1[3 * j + m]) || !isfinite(A.p1[3 * j + m])) bad = 1.0;
      if (block_sum(S, bad) != 0.0) {
        status = GX_NONFINITE;
        break;
      }
      bool keep = true;
      double next_tau1;
      if (PID) {
        const double scaled = sqrt(sq / (6.0 * (double)n));
        keep = scaled < 1.0;
        if (A.dtmin > 0.0) keep = keep || at_dtmin;
        const double inv = 1.0 / scaled;
        double f = (A.c1 == 0.0) ? 1.0 : pow(inv, A.c1);
        if (A.c2 != 0.0) f *= pow(prev_inv, A.c2);
        if (A.c3 != 0.0) f *= pow(prev_prev_inv, A.c3);
        const double fmin_ = keep ? 1.0 : A.factormin;
        const double factor = fmin(fmax(A.safety * f, fmin_), A.factormax);
        double dt = fmin(h * factor, A.dtmax);
        if (A.dtmin > 0.0) {
          if (!A.force_dtmin && dt < A.dtmin) status = GX_DT_MIN_REACHED;
          at_dtmin = dt <= A.dtmin;
          dt = fmax(dt, A.dtmin);
        }
        if (keep) {
          prev_prev_inv = prev_inv;
          prev_inv = (inv == 0.0 || isinf(inv)) ? 1.0 : inv;
        }
        next_tau1 = (keep ? tnext : tau) + dt;
      } else {
        next_tau1 = tnext + A.dt0;
      }
      if (keep) {
        // saves in (tau, tnext]: the step end, or the step's dense output
        // (diffrax's interpolant), which reads the end stage a(q1, tnext)
        bool have_end = M::SF == M::S;  // Dopri5's last stage is at q1
        while (si < A.T && dir * A.ts[si] <= tnext) {
          const double ts_tau = dir * A.ts[si];
          if (ts_tau == tnext) {
            nb_write(A, si, dir, A.q1, A.p1);
          } else {
            if (!have_end) {
              for (long long j = tid; j < n; j += kNbThreads) {
                S.x[j] = A.q1[3 * j];
                S.y[j] = A.q1[3 * j + 1];
                S.z[j] = A.q1[3 * j + 2];
              }
              __syncthreads();
              for (long long j = tid; j < n; j += kNbThreads) {
                double* g = A.G + ((long long)(M::SF - 1) * n + j) * 3;
                nb_accel(P, S, n, A.eps2, j, S.x[j], S.y[j], S.z[j], dir * tnext, g[0], g[1], g[2]);
              }
              __syncthreads();
              have_end = true;
            }
            double wq[M::SF], wp[M::SF];
            const double th = (ts_tau - tau) / h;
            dense_weights<M>(th, wq, wp);
            const double ht = h * th, hht = h * ht;
            for (long long j = tid; j < n; j += kNbThreads) {
              double sq_[3] = {0.0, 0.0, 0.0}, sp_[3] = {0.0, 0.0, 0.0};
              for (int l = 0; l < M::SF; ++l)
                for (int m = 0; m < 3; ++m) {
                  const double gl = A.G[((long long)l * n + j) * 3 + m];
                  sq_[m] = fma(wq[l], gl, sq_[m]);
                  sp_[m] = fma(wp[l], gl, sp_[m]);
                }
              double* o = A.out + (j * A.T + si) * 6;
              for (int m = 0; m < 3; ++m) {
                const double q0 = A.q[3 * j + m], p0 = A.p[3 * j + m];
                o[m] = fma(hht, sq_[m], fma(ht, p0, q0));
                o[3 + m] = dir * fma(ht, sp_[m], p0);
              }
            }
          }
          ++si;
        }
        __syncthreads();
        for (long long j = tid; j < n; j += kNbThreads)
          for (int m = 0; m < 3; ++m) {
            A.q[3 * j + m] = A.q1[3 * j + m];
            A.p[3 * j + m] = A.p1[3 * j + m];
          }
        ++nacc;
      }
      __syncthreads();
      const double tprev = fmin(keep ? tnext : tau, tau1);
      tnext = (next_tau1 > tau1 - 1e-10) ? (keep ? tau1 : tprev + 0.5 * (tau1 - tprev)) : next_tau1;
      tau = tprev;
      if (status == GX_DT_MIN_REACHED) break;
    }
  }
  if (status != GX_SUCCESS) {
    const double nanv = __longlong_as_double(0x7ff8000000000000ULL);
    for (; si < A.T; ++si)
      for (long long j = tid; j < n; j += kNbThreads)
        for (int m = 0; m < 6; ++m) A.out[(j * A.T + si) * 6 + m] = nanv;
  }
  if (tid == 0) {
    if (A.status) *A.status = status;
    if (A.nsteps) *A.nsteps = nst;
    if (A.naccept) *A.naccept = nacc;
  }
}

// The field alone (NBodyField.__call__'s acceleration): one thread per body,
// positions staged through LDS in tiles.
__global__ void __launch_bounds__(256) k_nbody_accel(DPot P, const double* __restrict__ xyz,
                                                     const double* __restrict__ gm, double eps2,
                                                     double t, long long n,
                                                     double* __restrict__ acc) {
  __shared__ double tx[256], ty[256], tz[256], tg[256];
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool own = j < n;
  const double x = own ? xyz[3 * j] : 0.0, y = own ? xyz[3 * j + 1] : 0.0,
               z = own ? xyz[3 * j + 2] : 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (own) accel<true>(P, x, y, z, t, ax, ay, az);
  for (long long k0 = 0; k0 < n; k0 += 256) {
    const long long k = k0 + threadIdx.x;
    __syncthreads();
    tx[threadIdx.x] = k < n ? xyz[3 * k] : 0.0;
    ty[threadIdx.x] = k < n ? xyz[3 * k + 1] : 0.0;
    tz[threadIdx.x] = k < n ? xyz[3 * k + 2] : 0.0;
    tg[threadIdx.x] = k < n ? gm[k] : 0.0;
    __syncthreads();
    const int m = (n - k0) < 256 ? (int)(n - k0) : 256;
    for (int i = 0; i < m; ++i) {
      const double dx = tx[i] - x, dy = ty[i] - y, dz = tz[i] - z;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double ir = rsqrt_fast(r2);
      const double w = (k0 + i == j) ? 0.0 : tg[i] * ir * ir * ir;
      ax = fma(w, dx, ax);
      ay = fma(w, dy, ay);
      az = fma(w, dz, az);
    }
  }
  if (own) {
    acc[3 * j] = ax;
    acc[3 * j + 1] = ay;
    acc[3 * j + 2] = az;
  }
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int64_t gx_nbody_workspace_bytes(int64_t n) {
  return n < 0 ? 0 : 256 + 8 * n * (12 + 3 * 14);
}

int gx_nbody_acceleration(const gx_pot* ext, const double* xyz, const double* gm,
                          double eps, double t, int64_t n, double* acc, void* stream) {
  if (!ext || !xyz || !gm || !acc || n < 0 || !(eps >= 0.0)) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if ((n + 255) / 256 > 0x7fffffffLL) return GX_EINVAL;
  hipLaunchKernelGGL(k_nbody_accel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, ext->dev, xyz, gm, eps * eps, t, (long long)n, acc);
  return check(hipGetLastError());
}

int gx_nbody_integrate(const gx_pot* ext, const double* gm, double eps,
                       const gx_solver_cfg* cfg, const double* w0, double t0, double t1,
                       const double* ts, int32_t T, int64_t n, double* out, int32_t* status,
                       int64_t* nsteps, int64_t* naccept, void* workspace, void* stream) {
  if (!ext || !gm || !cfg || !w0 || !ts || !out || !workspace) return GX_EINVAL;
  if (n < 1 || n > kNbMax || T < 1 || !(eps >= 0.0)) return GX_EINVAL;
  if (!std::isfinite(t0) || !std::isfinite(t1)) return GX_EINVAL;
  if (cfg->method != GX_DOPRI8 && cfg->method != GX_DOPRI5) return GX_EINVAL;
  if (cfg->controller != GX_CONSTANT && cfg->controller != GX_PID) return GX_EINVAL;
  if (cfg->controller == GX_CONSTANT && !(cfg->dt0 > 0.0)) return GX_EINVAL;
  if (cfg->controller == GX_PID &&
      !(cfg->rtol >= 0.0 && cfg->atol >= 0.0 && cfg->rtol + cfg->atol > 0.0 &&
        cfg->safety > 0.0 && cfg->factormin > 0.0 && cfg->factormax >= 1.0))
    return GX_EINVAL;
  NArgs A{};
  A.dt0 = cfg->dt0;
  A.rtol = cfg->rtol;
  A.atol = cfg->atol;
  A.safety = cfg->safety;
  A.factormin = cfg->factormin;
  A.factormax = cfg->factormax;
  A.dtmin = cfg->dtmin > 0.0 ? cfg->dtmin : 0.0;
  A.dtmax = cfg->dtmax > 0.0 ? cfg->dtmax : INFINITY;
  const double korder = cfg->method == GX_DOPRI8 ? 8.0 : 5.0;
  A.c1 = (cfg->icoeff + cfg->pcoeff + cfg->dcoeff) / korder;
  A.c2 = -(cfg->pcoeff + 2.0 * cfg->dcoeff) / korder;
  A.c3 = cfg->dcoeff / korder;
  A.max_steps = (cfg->max_steps > 0 && cfg->max_steps < 10000000) ? cfg->max_steps : 10000000;
  A.force_dtmin = cfg->force_dtmin;
  A.T = T;
  A.eps2 = eps * eps;
  A.t0 = t0;
  A.t1 = t1;
  A.gm = gm;
  A.w0 = w0;
  A.ts = ts;
  A.n = n;
  A.out = out;
  A.status = status;
  A.nsteps = reinterpret_cast<long long*>(nsteps);
  A.naccept = reinterpret_cast<long long*>(naccept);
  double* w = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256);
  A.q = w;
  A.p = w + 3 * n;
  A.q1 = w + 6 * n;
  A.p1 = w + 9 * n;
  A.G = w + 12 * n;
  hipStream_t s = (hipStream_t)stream;
  const bool pid = cfg->controller == GX_PID;
  if (cfg->method == GX_DOPRI8) {
    if (pid)
      hipLaunchKernelGGL((k_nbody<Dopri8, true>), dim3(1), dim3(kNbThreads), 0, s, ext->dev, A);
    else
      hipLaunchKernelGGL((k_nbody<Dopri8, false>), dim3(1), dim3(kNbThreads), 0, s, ext->dev, A);
  } else {
    if (pid)
      hipLaunchKernelGGL((k_nbody<Dopri5, true>), dim3(1), dim3(kNbThreads), 0, s, ext->dev, A);
    else
      hipLaunchKernelGGL((k_nbody<Dopri5, false>), dim3(1), dim3(kNbThreads), 0, s, ext->dev, A);
  }
  return check(hipGetLastError());
}

}  // extern "C"
