// gx_adjoint.hip -- discrete adjoint of fixed-step orbit integration.
//
// Gradients of a loss through n independent fixed-step Dopri8/Dopri5 orbit
// integrations, with respect to each orbit's initial state and the
// potential's table parameters.  galax differentiates its solves with
// jax (ForwardMode adjoint, dynamics/_src/orbit/solver.py:133; the stream
// loss derivatives of tests/functional/test_mockstreamgenerator.py:24-116);
// here the derivative of the discrete map is formed exactly, in reverse.
//
// Per orbit (one lane): a forward sweep stores the state every K steps; the
// reverse sweep walks the checkpoint segments backwards, recomputes each
// segment's states, and for every step recomputes its stage positions and
// applies the adjoint of the step.  In Runge-Kutta-Nystrom form (stages Q_i,
// accelerations g_i = a(Q_i), weights b, A, (bA), (A A)) the adjoint of one
// step of size h with output cotangents (qb', pb') is
//   gb_j = h b_j pb' + h^2 (bA)_j qb'                 (all stages j)
//   qb = qb',  pb = pb' + h (sum b) qb'
//   for k = S-1 .. 0:
//     Qb_k = -H(Q_k) gb_k          (Hessian-vector product, a = -grad Phi)
//     thb += (da/dtheta (Q_k))^T gb_k
//     qb += Qb_k;  pb += h rs_k Qb_k
//     gb_j += h^2 (A A)_kj Qb_k     (j < k)
// which is the transpose of the forward recurrence in gx_integrate.hip.
// A save time inside a step is the step's dense output,
//   q_s = q + th h p + h^2 th sum_l Wq_l(th) g_l,  p_s = p + h th sum_l Wp_l(th) g_l
// over the stages and the end stage g_E = a(q1) (Dopri8), as the forward
// pass produced it; its cotangents (cq, cp) add
//   gb_l += h^2 th Wq_l cq + h th Wp_l cp,  qb += cq,  pb += th h cq + cp
// and the end stage is reversed first: Q1b = -H(q1) gb_E, thb += ...,
// qb += Q1b, pb += h (sum b) Q1b, gb_j += h^2 (bA)_j Q1b.
#include <cmath>
#include <utility>

#include "gx_internal.hpp"
#include "gx_rk.hpp"
#include "gx_tableau.hpp"

namespace gx {
namespace {

constexpr int kAdjBlock = 128;
constexpr int kMaxTheta = 12;  // 3 table parameters x up to 4 components

struct AArgs {
  double dt0;
  const double* w0;
  const double* t0;
  const double* t1;
  int t_stride;
  int T;
  int ts_mode;
  int K;            // checkpoint interval (steps)
  long long C;      // checkpoint slots per orbit
  const double* ts;
  const double* cot;  // n x T x 6
  long long n;
  double* w0_bar;     // n x 6
  double* theta_bar;  // n x ntheta
  int ntheta;
  double* ckpt;  // [C][7][n]
  double* seg;   // [K][7][n]
  double* stq;   // [SF][3][n]: stage positions and the end position q1
  int* status;
};

__device__ __forceinline__ double save_time_a(const AArgs& A, long long i,
                                              int s, double t0, double t1) {
  switch (A.ts_mode) {
    case GX_TS_SHARED:
      return A.ts[s];
    case GX_TS_PER_ORBIT:
      return A.ts[i * A.T + s];
    case GX_TS_LINSPACE:
      return (s == A.T - 1) ? t1 : t0 + s * ((t1 - t0) / (A.T - 1));
    default:
      return t1;
  }
}

__device__ __forceinline__ double clip_end(double tprev, double tnext,
                                           double t1) {
  return (tnext > t1 - 1e-10) ? t1 : tnext;
}

// Forward step (no error estimate); optionally records the stage positions.
template <class M, class E, bool REC>
__device__ __forceinline__ void step_fwd(const DPot& P, const AArgs& A,
                                         long long i, double h, double (&q)[3],
                                         double (&p)[3]) {
  constexpr int S = M::S;
  double g[S][3];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    double Q[3];
    if (k == 0) {
      Q[0] = q[0];
      Q[1] = q[1];
      Q[2] = q[2];
    } else {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int l = 0; l < k; ++l) {
        const double a = M::k.aa[k][l];
        if (a != 0.0) {
          s0 = fma(a, g[l][0], s0);
          s1 = fma(a, g[l][1], s1);
          s2 = fma(a, g[l][2], s2);
        }
      }
      const double rs = M::k.rs[k];
      Q[0] = fma(h, fma(rs, p[0], h * s0), q[0]);
      Q[1] = fma(h, fma(rs, p[1], h * s1), q[1]);
      Q[2] = fma(h, fma(rs, p[2], h * s2), q[2]);
    }
    if (REC) {
#pragma unroll
      for (int m = 0; m < 3; ++m) A.stq[((long long)k * 3 + m) * A.n + i] = Q[m];
    }
    E::acc(P, Q[0], Q[1], Q[2], 0.0, g[k][0], g[k][1], g[k][2]);
  }
  double b0 = 0, b1 = 0, b2 = 0, c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
  for (int l = 0; l < S; ++l) {
    if (M::k.b[l] != 0.0) {
      b0 = fma(M::k.b[l], g[l][0], b0);
      b1 = fma(M::k.b[l], g[l][1], b1);
      b2 = fma(M::k.b[l], g[l][2], b2);
    }
    if (M::k.ba[l] != 0.0) {
      c0 = fma(M::k.ba[l], g[l][0], c0);
      c1 = fma(M::k.ba[l], g[l][1], c1);
      c2 = fma(M::k.ba[l], g[l][2], c2);
    }
  }
  constexpr double sb = M::k.sb;
  const double q0 = fma(h, fma(sb, p[0], h * c0), q[0]);
  const double q1 = fma(h, fma(sb, p[1], h * c1), q[1]);
  const double q2 = fma(h, fma(sb, p[2], h * c2), q[2]);
  if (REC && M::SF > S) {
    A.stq[((long long)S * 3 + 0) * A.n + i] = q0;
    A.stq[((long long)S * 3 + 1) * A.n + i] = q1;
    A.stq[((long long)S * 3 + 2) * A.n + i] = q2;
  }
  p[0] = fma(h, b0, p[0]);
  p[1] = fma(h, b1, p[1]);
  p[2] = fma(h, b2, p[2]);
  q[0] = q0;
  q[1] = q1;
  q[2] = q2;
}

// Sum over components of H w and (da/dtheta)^T w at Q.
template <int... K, size_t... I>
__device__ __forceinline__ void adj_all(const DPot& P, const double (&Q)[3],
                                        const double (&w)[3], double (&Hw)[3],
                                        double (&th)[kMaxTheta],
                                        std::index_sequence<I...>) {
  Hw[0] = Hw[1] = Hw[2] = 0.0;
  (comp_adjoint<K>(P.comps[I], Q[0], Q[1], Q[2], w, Hw, th + 3 * I), ...);
}

// Adjoint of one step of size h from (q, p) given the step end's output
// cotangents (qb, pb), updated in place to the cotangents of (q, p).  The
// save times inside the step add gx (their stage cotangents over the SF
// slots) and dq, dp (their direct cotangents of q and p); with no such save
// (any_save false) those are zero and the end stage is skipped.
template <class M, class E, int... K>
__device__ __forceinline__ void step_adj(const DPot& P, const AArgs& A,
                                         long long i, double h,
                                         const double (&q)[3],
                                         const double (&p)[3], double (&qb)[3],
                                         double (&pb)[3],
                                         double (&th)[kMaxTheta],
                                         const double (&gx)[M::SF][3],
                                         const double (&dq)[3],
                                         const double (&dp)[3],
                                         bool any_save) {
  constexpr int S = M::S;
  constexpr int SF = M::SF;
  {
    double qq[3] = {q[0], q[1], q[2]}, pp[3] = {p[0], p[1], p[2]};
    step_fwd<M, E, true>(P, A, i, h, qq, pp);
  }
  double gb[SF][3];
  const double h2 = h * h;
#pragma unroll
  for (int j = 0; j < SF; ++j)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      gb[j][m] = (j < S ? fma(h * M::k.b[j], pb[m], h2 * M::k.ba[j] * qb[m]) : 0.0) + gx[j][m];
  constexpr double sb = M::k.sb;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    pb[m] = fma(h * sb, qb[m], pb[m]) + dp[m];
    qb[m] += dq[m];
  }
  if constexpr (SF > S) {
    if (any_save) {
      // the end stage g_E = a(q1), q1 = q + h (sb p + h sum (bA)_j g_j)
      double Q[3], Hw[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) Q[m] = A.stq[((long long)S * 3 + m) * A.n + i];
      adj_all<K...>(P, Q, gb[S], Hw, th, std::make_index_sequence<sizeof...(K)>{});
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        qb[m] -= Hw[m];
        pb[m] = fma(-h * sb, Hw[m], pb[m]);
      }
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double a = M::k.ba[j];
        if (a != 0.0) {
#pragma unroll
          for (int m = 0; m < 3; ++m) gb[j][m] = fma(-h2 * a, Hw[m], gb[j][m]);
        }
      }
    }
  }
#pragma unroll
  for (int k = S - 1; k >= 0; --k) {
    double Q[3], Hw[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) Q[m] = A.stq[((long long)k * 3 + m) * A.n + i];
    adj_all<K...>(P, Q, gb[k], Hw, th, std::make_index_sequence<sizeof...(K)>{});
    // Qb_k = -H gb_k
    const double rs = M::k.rs[k];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      qb[m] -= Hw[m];
      pb[m] = fma(-h * rs, Hw[m], pb[m]);
    }
#pragma unroll
    for (int j = 0; j < k; ++j) {
      const double a = M::k.aa[k][j];
      if (a != 0.0) {
#pragma unroll
        for (int m = 0; m < 3; ++m) gb[j][m] = fma(-h2 * a, Hw[m], gb[j][m]);
      }
    }
  }
}

template <class M, class E, int... K>
__global__ void __launch_bounds__(kAdjBlock) k_adjoint(DPot P, AArgs A) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double t0v = A.t0[A.t_stride ? i : 0];
  const double t1v = A.t1[A.t_stride ? i : 0];
  const double dir = (t0v < t1v) ? 1.0 : -1.0;
  const double tau0 = dir * t0v, tau1 = dir * t1v;
  const double* w = A.w0 + i * 6;
  double q[3] = {w[0], w[1], w[2]};
  double p[3] = {dir * w[3], dir * w[4], dir * w[5]};
  double th[kMaxTheta];
#pragma unroll
  for (int k = 0; k < kMaxTheta; ++k) th[k] = 0.0;

  // ---- forward sweep with checkpoints every K steps
  long long N = 0;
  double tau = tau0;
  int status = GX_SUCCESS;
  if (t0v != t1v) {
    double tn = fmin(tau + A.dt0, tau1);
    while (tau < tau1) {
      if (N % A.K == 0) {
        const long long c = N / A.K;
        if (c >= A.C) {
          status = GX_MAX_STEPS_REACHED;
          break;
        }
        double* ck = A.ckpt + c * 7 * A.n + i;
        ck[0 * A.n] = q[0];
        ck[1 * A.n] = q[1];
        ck[2 * A.n] = q[2];
        ck[3 * A.n] = p[0];
        ck[4 * A.n] = p[1];
        ck[5 * A.n] = p[2];
        ck[6 * A.n] = tau;
      }
      step_fwd<M, E, false>(P, A, i, tn - tau, q, p);
      ++N;
      tau = tn;
      tn = clip_end(tau, tau + A.dt0, tau1);
    }
  }
  // ---- reverse sweep
  double qb[3] = {0.0, 0.0, 0.0}, pb[3] = {0.0, 0.0, 0.0};
  int si = A.T - 1;
  auto cot = [&](int s, double (&cq)[3], double (&cp)[3]) {
    const double* c = A.cot + (i * A.T + s) * 6;
    cq[0] = c[0];
    cq[1] = c[1];
    cq[2] = c[2];
    cp[0] = dir * c[3];
    cp[1] = dir * c[4];
    cp[2] = dir * c[5];
  };
  if (status == GX_SUCCESS) {
    const long long nseg = (N + A.K - 1) / A.K;
    for (long long c = nseg - 1; c >= 0; --c) {
      // recompute the segment's step starts
      const long long n0 = c * A.K;
      const int len = (int)((N - n0) < A.K ? (N - n0) : A.K);
      double qs[3], ps[3];
      const double* ck = A.ckpt + c * 7 * A.n + i;
      for (int m = 0; m < 3; ++m) {
        qs[m] = ck[m * A.n];
        ps[m] = ck[(m + 3) * A.n];
      }
      double ts_ = ck[6 * A.n];
      for (int m = 0; m < len; ++m) {
        double* sg = A.seg + (long long)m * 7 * A.n + i;
        sg[0 * A.n] = qs[0];
        sg[1 * A.n] = qs[1];
        sg[2 * A.n] = qs[2];
        sg[3 * A.n] = ps[0];
        sg[4 * A.n] = ps[1];
        sg[5 * A.n] = ps[2];
        sg[6 * A.n] = ts_;
        if (m + 1 < len) {
          const double tn = (n0 + m == 0) ? fmin(ts_ + A.dt0, tau1)
                                          : clip_end(ts_, ts_ + A.dt0, tau1);
          step_fwd<M, E, false>(P, A, i, tn - ts_, qs, ps);
          ts_ = tn;
        }
      }
      // reverse over the segment's steps
      for (int m = len - 1; m >= 0; --m) {
        const double* sg = A.seg + (long long)m * 7 * A.n + i;
        const double qn[3] = {sg[0], sg[A.n], sg[2 * A.n]};
        const double pn[3] = {sg[3 * A.n], sg[4 * A.n], sg[5 * A.n]};
        const double tau_n = sg[6 * A.n];
        const long long n = n0 + m;
        const double tau_np1 = (n == 0) ? fmin(tau_n + A.dt0, tau1)
                                        : clip_end(tau_n, tau_n + A.dt0, tau1);
        // saves at the step end
        while (si >= 0 && dir * save_time_a(A, i, si, t0v, t1v) == tau_np1) {
          double cq[3], cp[3];
          cot(si, cq, cp);
          for (int k = 0; k < 3; ++k) {
            qb[k] += cq[k];
            pb[k] += cp[k];
          }
          --si;
        }
        // interior saves: the step's dense output at th = (t_s - tau_n) / h
        const double h = tau_np1 - tau_n;
        double gx[M::SF][3], dq[3] = {0.0, 0.0, 0.0}, dp[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int l = 0; l < M::SF; ++l) gx[l][0] = gx[l][1] = gx[l][2] = 0.0;
        bool any_save = false;
        while (si >= 0) {
          const double tsv = dir * save_time_a(A, i, si, t0v, t1v);
          if (!(tsv > tau_n && tsv < tau_np1)) break;
          double cq[3], cp[3];
          cot(si, cq, cp);
          const double tht = (tsv - tau_n) / h;
          double wq[M::SF], wp[M::SF];
          dense_weights<M>(tht, wq, wp);
          const double ht = h * tht, hht = h * ht;
#pragma unroll
          for (int l = 0; l < M::SF; ++l)
#pragma unroll
            for (int k = 0; k < 3; ++k)
              gx[l][k] = fma(hht * wq[l], cq[k], fma(ht * wp[l], cp[k], gx[l][k]));
          for (int k = 0; k < 3; ++k) {
            dq[k] += cq[k];
            dp[k] = fma(ht, cq[k], dp[k] + cp[k]);
          }
          any_save = true;
          --si;
        }
        step_adj<M, E, K...>(P, A, i, h, qn, pn, qb, pb, th, gx, dq, dp, any_save);
      }
    }
    // saves at (or before) the start
    while (si >= 0) {
      double cq[3], cp[3];
      cot(si, cq, cp);
      for (int k = 0; k < 3; ++k) {
        qb[k] += cq[k];
        pb[k] += cp[k];
      }
      --si;
    }
  }
  double* wb = A.w0_bar + i * 6;
  wb[0] = qb[0];
  wb[1] = qb[1];
  wb[2] = qb[2];
  wb[3] = dir * pb[0];
  wb[4] = dir * pb[1];
  wb[5] = dir * pb[2];
  for (int k = 0; k < A.ntheta && k < kMaxTheta; ++k)
    A.theta_bar[i * A.ntheta + k] = th[k];
  if (A.status) A.status[i] = status;
}

template <class M, int... K>
int launch_adj(const gx_pot* pot, AArgs& A, hipStream_t s) {
  using E = EvalFixed<K...>;
  const long long blocks = (A.n + kAdjBlock - 1) / kAdjBlock;
  hipLaunchKernelGGL((k_adjoint<M, E, K...>), dim3((unsigned)blocks),
                     dim3(kAdjBlock), 0, s, pot->dev, A);
  return check(hipGetLastError());
}

template <class M>
int dispatch_adj(const gx_pot* pot, AArgs& A, hipStream_t s) {
  if (pot->sig == kSigMilkyWay)
    return launch_adj<M, GX_MIYAMOTO_NAGAI, GX_NFW, GX_HERNQUIST, GX_HERNQUIST>(pot, A, s);
  if (pot->ncomp == 1 && pot->basic) {
    switch (pot->kind0) {
      case GX_KEPLER:
        return launch_adj<M, GX_KEPLER>(pot, A, s);
      case GX_HERNQUIST:
        return launch_adj<M, GX_HERNQUIST>(pot, A, s);
      case GX_MIYAMOTO_NAGAI:
        return launch_adj<M, GX_MIYAMOTO_NAGAI>(pot, A, s);
      case GX_NFW:
        return launch_adj<M, GX_NFW>(pot, A, s);
      case GX_PLUMMER:
        return launch_adj<M, GX_PLUMMER>(pot, A, s);
      default:
        break;
    }
  }
  return GX_EINVAL;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int64_t gx_adjoint_workspace_bytes(int64_t n, int64_t max_steps, int32_t K) {
  if (n < 0 || max_steps < 0 || K < 1) return 0;
  const int64_t C = max_steps / K + 2;
  return 8 * n * (7 * C + 7 * (int64_t)K + 3 * 14) + 256;
}

int gx_adjoint(const gx_pot* pot, const gx_solver_cfg* cfg, const double* w0,
               const double* t0, const double* t1, int32_t t_stride,
               const double* ts, int32_t T, int32_t ts_mode, int64_t n,
               const double* cot, double* w0_bar, double* theta_bar,
               int32_t ntheta, int64_t max_steps, int32_t K, int32_t* status,
               void* workspace, void* stream) {
  if (!pot || !cfg || n < 0 || T < 1 || K < 1 || max_steps < 1) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!w0 || !t0 || !t1 || !cot || !w0_bar || !workspace) return GX_EINVAL;
  if (ntheta < 0 || ntheta > kMaxTheta || (ntheta > 0 && !theta_bar)) return GX_EINVAL;
  if (cfg->controller != GX_CONSTANT || !(cfg->dt0 > 0.0)) return GX_EINVAL;
  if (cfg->method != GX_DOPRI8 && cfg->method != GX_DOPRI5) return GX_EINVAL;
  if (t_stride != 0 && t_stride != 1) return GX_EINVAL;
  if (ts_mode < GX_TS_SHARED || ts_mode > GX_TS_FINAL) return GX_EINVAL;
  if ((ts_mode == GX_TS_SHARED || ts_mode == GX_TS_PER_ORBIT) && !ts) return GX_EINVAL;
  if (ts_mode == GX_TS_FINAL && T != 1) return GX_EINVAL;
  AArgs A{};
  A.dt0 = cfg->dt0;
  A.w0 = w0;
  A.t0 = t0;
  A.t1 = t1;
  A.t_stride = t_stride;
  A.T = T;
  A.ts_mode = ts_mode;
  A.K = K;
  A.C = max_steps / K + 2;
  A.ts = ts;
  A.cot = cot;
  A.n = n;
  A.w0_bar = w0_bar;
  A.theta_bar = theta_bar;
  A.ntheta = ntheta;
  A.status = status;
  char* ws = static_cast<char*>(workspace);
  A.ckpt = reinterpret_cast<double*>(ws);
  A.seg = A.ckpt + 7 * A.C * n;
  A.stq = A.seg + 7 * (long long)K * n;
  hipStream_t s = (hipStream_t)stream;
  if (cfg->method == GX_DOPRI8) return dispatch_adj<Dopri8>(pot, A, s);
  return dispatch_adj<Dopri5>(pot, A, s);
}

}  // extern "C"
