// gx_integrate.hip -- the orbit integrator of libgalaxhip.so.
//
// Replaces the diffrax.diffeqsolve loop that OrbitSolver.solve runs for
// gd.evaluate_orbit (src/galax/dynamics/_src/orbit/solver.py:774-1039,
// legacy/funcs.py:42-213), with the batched-t0/t1 semantics: every orbit has
// its own step sequence.
//
// Design (one orbit per lane, persistent launch):
//   * the grid is sized to the number of resident waves; lanes take orbits
//     as tickets from a device-wide atomic counter (the work queue), so lanes
//     never idle while orbits of uneven length remain;
//   * the head of a cost-ordered queue (the longest orbits) bounds the
//     launch.  A wave stalls as a whole whenever one of its lanes takes a
//     ticket (a device-scope atomic, then dependent loads of the queue
//     entry and the new orbit's state), so in a large grid the first
//     kHeadLanes lanes integrate tickets 0..kHeadLanes-1 and take no more:
//     their waves run the bare step loop at top wave priority, and the rest
//     of the grid drains the queue;
//   * the state, the stage accelerations (Runge-Kutta-Nystrom form, 3
//     doubles per stage), the error norm and the PID controller live in
//     registers; all arithmetic is FP64;
//   * each loop trip is one attempted step.  Its last stage sits at the
//     step's end (first same as last), so an accepted step hands the next
//     one its first stage, and the save times inside an accepted step are
//     served by the method's dense output from the step's own stages
//     (diffrax's SaveAt(ts) interpolant), not by further field evaluations;
//   * integration backwards in time runs in tau = -t with the velocity sign
//     flipped, as diffrax does.
// The controller follows diffrax 0.7.0's PIDController (error_order = method
// order, RMS norm, factormin applied to rejected steps only), its automatic
// initial step and its clip-to-end rule.
#include <cmath>
#include <mutex>

#include "gx_internal.hpp"
#include "gx_rk.hpp"

namespace gx {
namespace {

struct IArgs {
  double dt0, rtol, atol, safety, factormin, factormax, dtmin, dtmax;
  double c1, c2, c3;  // PID exponents (i+p+d)/k, -(p+2d)/k, d/k
  long long max_steps;
  int force_dtmin;
  int t_stride;
  int T;
  int ts_mode;
  int root8;  // c1 == 1/8 and c2 == c3 == 0: factor from three square roots
  const double* w0;
  const double* t0;
  const double* t1;
  const double* ts;
  long long n;
  const long long* order;
  double* out;
  int* status;
  long long* nsteps;
  long long* naccept;
  unsigned long long* counter;
  double* h0;  // first step per orbit, from the prepass (workspace)
  double* a0;  // start acceleration per orbit [n][3], from the prepass
  long long head_lanes;  // lanes that start on a fixed ticket
  int head_only_once;    // head lanes take no second ticket
  int chunk;             // tickets taken from the queue at a time
  double* rec;           // dense records [n][rec_cap][kDenseRec] (dense mode)
  long long rec_cap;
  long long* nrec;       // records written per orbit
};

constexpr int kBlock = 256;
constexpr long long kStepCap = 10000000;  // attempted steps per orbit
// The start of an orbit in solver variables: its acceleration a0 = a(q, tau0)
// (the first step's first stage) and the step the controller starts with:
// dt0 if given; dtmin if the controller sets one (the reference's
// integrate_field doctest with dtmin = 0.05, dynamics/_src/solver.py:342-365,
// reaches its printed values only from that first step); otherwise diffrax's
// automatic choice (Hairer, Norsett & Wanner II.4; final exponent 1/(p+1)
// for a method of order p); then clipped to [dtmin, dtmax].  One call site
// of the field, run once or (automatic rule) twice.  Quotients, square
// roots and the final root use the short primitives (each within about an
// ulp of the correctly rounded operation): a lane that starts an orbit runs
// this while the rest of its wave waits.
template <class E, int ORDER, bool PID>
__device__ double start_orbit(const DPot& P, const IArgs& A, double dir,
                              double tau0, const double (&q)[3],
                              const double (&p)[3], double (&a0)[3],
                              int& at_dtmin) {
  const bool automatic = !(A.dt0 > 0.0) && !(PID && A.dtmin > 0.0);
  double isc[6];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    isc[m] = rcp(A.atol + fabs(q[m]) * A.rtol);
    isc[m + 3] = rcp(A.atol + fabs(p[m]) * A.rtol);
  }
  double q1[3], p1[3];
  double d1 = 0.0, h0 = 0.0, d2 = 0.0;
  const int npass = automatic ? 2 : 1;
#pragma nounroll
  for (int pass = 0; pass < npass; ++pass) {
    double x = q[0], y = q[1], z = q[2], t = tau0;
    if (pass == 1) {
      x = q1[0];
      y = q1[1];
      z = q1[2];
      t = tau0 + h0;
    }
    double ax, ay, az;
    E::acc(P, x, y, z, dir * t, ax, ay, az);
    if (pass == 0) {
      a0[0] = ax;
      a0[1] = ay;
      a0[2] = az;
      if (automatic) {
        double d0 = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const double u = q[m] * isc[m], v = p[m] * isc[m + 3];
          d0 += u * u + v * v;
          const double fu = p[m] * isc[m], fv = a0[m] * isc[m + 3];
          d1 += fu * fu + fv * fv;
        }
        d0 = sqrt_fast(div6(d0));
        d1 = sqrt_fast(div6(d1));
        const bool cond = (d0 < 1e-5) || (d1 < 1e-5);
        h0 = cond ? 1e-6 : 0.01 * div(d0, cond ? 1.0 : d1);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          q1[m] = q[m] + p[m] * h0;
          p1[m] = p[m] + a0[m] * h0;
        }
      }
    } else {
      const double a1[3] = {ax, ay, az};
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double u = (p1[m] - p[m]) * isc[m],
                     v = (a1[m] - a0[m]) * isc[m + 3];
        d2 += u * u + v * v;
      }
      d2 = div(sqrt_fast(div6(d2)), h0);
    }
  }
  double h;
  if (automatic) {
    const double maxd = fmax(d1, d2);
    const double h1 = (maxd <= 1e-15) ? fmax(1e-6, h0 * 1e-3)
                                      : root_n<ORDER + 1>(div(0.01, maxd));
    h = fmin(100.0 * h0, h1);
  } else {
    h = (A.dt0 > 0.0) ? A.dt0 : A.dtmin;
  }
  at_dtmin = 0;
  if (PID) {
    h = fmin(h, A.dtmax);
    if (A.dtmin > 0.0) {
      at_dtmin = h <= A.dtmin ? 1 : 0;
      h = fmax(h, A.dtmin);
    }
  }
  return h;
}

__device__ __forceinline__ double clip_to_end(double tprev, double tnext,
                                              double t1, bool keep) {
  return (tnext > t1 - 1e-10) ? (keep ? t1 : tprev + 0.5 * (t1 - tprev))
                              : tnext;
}

__device__ __forceinline__ double save_time(const IArgs& A, long long orb,
                                            int s, double t0, double t1) {
  switch (A.ts_mode) {
    case GX_TS_SHARED:
      return A.ts[s];
    case GX_TS_PER_ORBIT:
      return A.ts[orb * A.T + s];
    case GX_TS_LINSPACE:
      return (s == A.T - 1) ? t1 : t0 + s * ((t1 - t0) / (A.T - 1));
    default:
      return t1;
  }
}

__device__ __forceinline__ void write_save(const IArgs& A, long long orb,
                                           int s, double dir,
                                           const double (&q)[3],
                                           const double (&p)[3]) {
  double* o = A.out + (orb * A.T + s) * 6;
  o[0] = q[0];
  o[1] = q[1];
  o[2] = q[2];
  o[3] = dir * p[0];
  o[4] = dir * p[1];
  o[5] = dir * p[2];
}

__device__ __forceinline__ long long ticket_to_orbit(const IArgs& A,
                                                     long long idx) {
  if (idx >= A.n) return -1;
  if (!A.order) return idx;
  const long long o = A.order[idx];
  return (o >= 0 && o < A.n) ? o : -1;  // the host validates permutations
}

// Start-of-orbit prepass for the evaluators that do not compute it inline:
// one lane per orbit writes the controller's first step h0[orbit] and the
// start acceleration a0[orbit] (the first stage of the first step), so that
// the persistent kernel carries no second copy of their field evaluation.
// The field is the integrator's own evaluator, so a0 equals what the
// persistent kernel would compute.  (The small Milky Way evaluator computes
// both in the persistent kernel instead: a separate launch would wait for
// resident slots behind concurrent integrator launches.)
template <int ORDER, bool PID, class E>
__global__ void __launch_bounds__(256) k_first_step(DPot P, IArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       i < A.n; i += stride) {
    const double t0v = A.t0[A.t_stride ? i : 0];
    const double t1v = A.t1[A.t_stride ? i : 0];
    if (t0v == t1v) continue;
    const double dir = (t0v < t1v) ? 1.0 : -1.0;
    const double* w = A.w0 + i * 6;
    const double q[3] = {w[0], w[1], w[2]};
    const double p[3] = {dir * w[3], dir * w[4], dir * w[5]};
    double a0[3];
    int at_dtmin;
    A.h0[i] = start_orbit<E, ORDER, PID>(P, A, dir, dir * t0v, q, p, a0,
                                          at_dtmin);
    A.a0[3 * i] = a0[0];
    A.a0[3 * i + 1] = a0[1];
    A.a0[3 * i + 2] = a0[2];
  }
}

// Per-lane state that the step itself does not touch lives in LDS slots
// (structure of arrays, one column per lane: conflict-free), which keeps the
// register file for the stage accelerations.
struct Cold {
  double tau1[kBlock];    // end time (solver time)
  double t0[kBlock], t1[kBlock];
  double prev_inv[kBlock], prev_prev_inv[kBlock];
  long long nst[kBlock], nacc[kBlock];
  long long qnext[kBlock], qend[kBlock];  // the lane's unused tickets
  int si[kBlock], st[kBlock], at_dtmin[kBlock];
};

// The lane's next ticket, or -1: the queue is empty, or the lane is a head
// lane of a grid large enough for the rest to drain the queue.  A lane
// takes `chunk` consecutive tickets per atomic and uses them in turn.
__device__ __forceinline__ long long next_orbit(const IArgs& A, Cold& C,
                                                int L, bool head) {
  if (head && A.head_only_once) return -1;
  long long idx = C.qnext[L];
  if (idx >= C.qend[L]) {
    idx = A.head_lanes +
          (long long)atomicAdd(A.counter, (unsigned long long)A.chunk);
    C.qend[L] = idx + A.chunk;
  }
  C.qnext[L] = idx + 1;
  return ticket_to_orbit(A, idx);
}

// Evaluators that stage their track in the launch's dynamic LDS
// (E::kLdsTrack, with E::stage).
template <class E, class = void>
struct lds_track : std::false_type {};
template <class E>
struct lds_track<E, std::void_t<decltype(E::kLdsTrack)>>
    : std::bool_constant<E::kLdsTrack> {};

template <class M, bool PID, class E, bool DENSE>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(E::kWavesPerSimd > 0 ? E::kWavesPerSimd : 1)))
    k_integrate(DPot P, IArgs A) {
  __shared__ Cold C;
  if constexpr (lds_track<E>::value) {
    E::stage(P);
    __syncthreads();
  }
  const int L = threadIdx.x;
  const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // The first tickets of a cost-ordered queue hold the longest orbits, whose
  // step latency bounds the launch: their lanes start on them, and their
  // waves issue ahead of any other wave on their SIMD.
  const bool head = lane < A.head_lanes;
  C.qnext[L] = 0;
  C.qend[L] = 0;
  long long orb = head ? ticket_to_orbit(A, lane) : next_orbit(A, C, L, false);
  bool fresh = true;
  if (head) __builtin_amdgcn_s_setprio(3);

  constexpr int SF = M::SF;
  // q, p at the current step start; g[0] = a(q, tau) carried from the last
  // accepted step's end stage, g[1..SF-1] the stages of the step in hand
  double q[3], p[3], g[SF][3];
  double tau = 0.0, tnext = 0.0, dir = 1.0;

  while (orb >= 0) {
    bool done = false;
    if (fresh) {
      fresh = false;
      const double t0v = A.t0[A.t_stride ? orb : 0];
      const double t1v = A.t1[A.t_stride ? orb : 0];
      const double* w = A.w0 + orb * 6;
      dir = (t0v < t1v) ? 1.0 : -1.0;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        q[m] = w[m];
        p[m] = dir * w[m + 3];
      }
      tau = dir * t0v;
      C.tau1[L] = dir * t1v;
      C.t0[L] = t0v;
      C.t1[L] = t1v;
      C.prev_inv[L] = 1.0;
      C.prev_prev_inv[L] = 1.0;
      C.nst[L] = 0;
      C.nacc[L] = 0;
      C.st[L] = GX_SUCCESS;
      int si = 0;
      if (t0v == t1v) {
        for (; si < A.T; ++si) write_save(A, orb, si, dir, q, p);
        done = true;
      } else {
        double h0;
        int at_dtmin;
        if constexpr (E::kInlineFirstStep) {
          h0 = start_orbit<E, M::order, PID>(P, A, dir, tau, q, p, g[0],
                                             at_dtmin);
        } else {
          h0 = A.h0[orb];
          g[0][0] = A.a0[3 * orb];
          g[0][1] = A.a0[3 * orb + 1];
          g[0][2] = A.a0[3 * orb + 2];
          at_dtmin = (PID && A.dtmin > 0.0 && h0 == A.dtmin) ? 1 : 0;
        }
        C.at_dtmin[L] = at_dtmin;
        tnext = fmin(tau + h0, dir * t1v);
        // saves at the start time
        while (si < A.T && dir * save_time(A, orb, si, t0v, t1v) <= tau) {
          write_save(A, orb, si, dir, q, p);
          ++si;
        }
      }
      C.si[L] = si;
    }
    if (!done) {
      if (C.nst[L] >= A.max_steps) {
        C.st[L] = GX_MAX_STEPS_REACHED;
        done = true;
      } else if (!(tnext - tau >= 0.0)) {
        C.st[L] = GX_NONFINITE;  // NaN step size: stop rather than spin
        done = true;
      } else {
        // ---- one attempted step
        const double h = tnext - tau;
        double q1[3], p1[3], eq[3], ep[3];
        rk_step<M, E, PID>(P, dir, tau, h, tnext, q, p, g, q1, p1, eq, ep);
        const double tau1 = C.tau1[L];
        C.nst[L] += 1;
        bool finite = true;
#pragma unroll
        for (int m = 0; m < 3; ++m)
          finite = finite && isfinite(q1[m]) && isfinite(p1[m]);
        if (!finite) {
          C.st[L] = GX_NONFINITE;
          done = true;
        } else {
          bool keep = true;
          double next_t1;
          if (PID) {
            // diffrax's RMS norm of the scaled error, here through its
            // mean square ms (the norm is below 1 exactly when ms is);
            // quotients and roots by the short primitives, within about
            // an ulp of the correctly rounded operations
            double sq = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              const double uq =
                  eq[m] * rcp(A.atol + fmax(fabs(q[m]), fabs(q1[m])) * A.rtol);
              const double up =
                  ep[m] * rcp(A.atol + fmax(fabs(p[m]), fabs(p1[m])) * A.rtol);
              sq += uq * uq + up * up;
            }
            const double ms = div6(sq);
            keep = ms < 1.0;
            const bool at_dtmin = C.at_dtmin[L] != 0;
            if (A.dtmin > 0.0) keep = keep || at_dtmin;
            // inv = 1 / norm = ms^(-1/2); zero (and subnormal) ms gives
            // +inf, infinite ms 0, NaN stays NaN
            const bool normal = ms >= kMinNormal && ms <= kMaxFinite;
            const double special = (ms < kMinNormal)   ? __builtin_inf()
                                   : (ms > kMaxFinite) ? 0.0
                                                       : ms;
            const double a = rsqrt_fast(ms);
            const double inv = normal ? a : special;
            double f;
            if (A.root8) {
              // inv^(1/8) = 1 / rsqrt(rsqrt(rsqrt(inv)))
              f = normal ? rcp(rsqrt_fast(rsqrt_fast(rsqrt_fast(a)))) : special;
            } else {
              f = (A.c1 == 0.0) ? 1.0 : pow(inv, A.c1);
              if (A.c2 != 0.0) f *= pow(C.prev_inv[L], A.c2);
              if (A.c3 != 0.0) f *= pow(C.prev_prev_inv[L], A.c3);
            }
            const double fmin_ = keep ? 1.0 : A.factormin;
            const double factor =
                fmin(fmax(A.safety * f, fmin_), A.factormax);
            double dt = fmin(h * factor, A.dtmax);
            if (A.dtmin > 0.0) {
              if (!A.force_dtmin && dt < A.dtmin)
                C.st[L] = GX_DT_MIN_REACHED;
              C.at_dtmin[L] = dt <= A.dtmin ? 1 : 0;
              dt = fmax(dt, A.dtmin);
            }
            if (keep) {
              // a zero or infinite error does not steer later steps
              C.prev_prev_inv[L] = C.prev_inv[L];
              C.prev_inv[L] = (inv == 0.0 || isinf(inv)) ? 1.0 : inv;
            }
            next_t1 = (keep ? tnext : tau) + dt;
          } else {
            next_t1 = tnext + A.dt0;
          }
          if (keep) {
            rk_end_stage<M, E>(P, dir, tnext, q1, g);
            // save times in (tau, tnext]: the step's end, or its dense
            // output at th = (t_s - tau) / h
            int si = C.si[L];
            const double t0v = C.t0[L], t1v = C.t1[L];
            while (si < A.T) {
              const double ts_tau = dir * save_time(A, orb, si, t0v, t1v);
              if (ts_tau > tnext) break;
              if (ts_tau == tnext) {
                write_save(A, orb, si, dir, q1, p1);
              } else {
                double qs[3], ps[3];
                dense_eval<M>(div(ts_tau - tau, h), h, q, p, g, qs, ps);
                write_save(A, orb, si, dir, qs, ps);
              }
              ++si;
            }
            C.si[L] = si;
            if constexpr (DENSE) {
              const long long k = C.nacc[L];
              if (k < A.rec_cap)
                write_record<M>(A.rec + (orb * A.rec_cap + k) * kDenseRec, dir,
                                tau, h, q, p, g);
              else
                C.st[L] = GX_MAX_STEPS_REACHED;  // out of record slots
            }
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              q[m] = q1[m];
              p[m] = p1[m];
              g[0][m] = g[SF - 1][m];
            }
            tau = tnext;
            C.nacc[L] += 1;
          }
          const double tprev = fmin(tau, tau1);
          tnext = clip_to_end(tprev, next_t1, tau1, keep);
          tau = tprev;
          if (C.st[L] == GX_DT_MIN_REACHED) done = true;
          if (DENSE && C.st[L] == GX_MAX_STEPS_REACHED) done = true;
          done = done || tau >= tau1;
        }
      }
    }
    if (done) {
      int si = C.si[L];
      const int st = C.st[L];
      if (st != GX_SUCCESS || si < A.T) {
        const double nanv = __longlong_as_double(0x7ff8000000000000ULL);
        const double nq[3] = {nanv, nanv, nanv};
        for (; si < A.T; ++si) write_save(A, orb, si, 1.0, nq, nq);
      }
      if (A.status) A.status[orb] = st;
      if (A.nsteps) A.nsteps[orb] = C.nst[L];
      if (A.naccept) A.naccept[orb] = C.nacc[L];
      if constexpr (DENSE)
        A.nrec[orb] = C.nacc[L] < A.rec_cap ? C.nacc[L] : A.rec_cap;
      orb = next_orbit(A, C, L, head);
      fresh = true;
    }
  }
}

// Clear the queue counter (a kernel, not a memset, so that a captured
// launch sequence replays correctly).
__global__ void k_reset_queue(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ULL;
}

template <class M, bool PID, class E, bool DENSE = false>
int launch(const gx_pot* pot, IArgs& A, hipStream_t stream) {
  const auto kern = k_integrate<M, PID, E, DENSE>;
  const LaunchShape& ls = launch_shape();
  const int block = kBlock;
  // dynamic LDS: the staged track of an LDS-track evaluator
  size_t dyn = 0;
  if constexpr (lds_track<E>::value) {
    dyn = sizeof(double) * (size_t)pot->dev.trk.lds_doubles;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)dyn) != hipSuccess)
      return GX_ELAUNCH;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, dyn) !=
          hipSuccess ||
      occ < 1)
    occ = 1;
  int per_cu = ls.blocks_per_cu;
  if (per_cu > occ) per_cu = occ;
  // one launch takes 1/parts of the resident capacity, so that `parts`
  // launches on different streams can run side by side
  long long blocks = (long long)num_cus(pot->device) * per_cu / ls.parts;
  const long long need = (A.n + block - 1) / block;
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  // Head lanes take one ticket only when the rest of the grid is large
  // enough to drain the queue (at least three times the head).
  const long long head = ls.head_lanes;
  A.head_lanes = head < blocks * block ? head : blocks * block;
  A.head_only_once = (head > 0 && blocks * block >= 4 * head) ? 1 : 0;
  A.chunk = ls.chunk;
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, stream, A.counter);
  if constexpr (!E::kInlineFirstStep) {
    long long fblocks = (A.n + 255) / 256;
    const long long fcap = (long long)num_cus(pot->device) * 8;
    if (fblocks > fcap) fblocks = fcap;
    hipLaunchKernelGGL((k_first_step<M::order, PID, E>),
                       dim3((unsigned)fblocks), dim3(256), 0, stream, pot->dev,
                       A);
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(block), dyn, stream,
                     pot->dev, A);
  return check(hipGetLastError());
}

// The dense solution of n orbits from their step records at query times:
// one lane per (orbit, time).  The step holding t is found by bisection on
// the records' start times (ordered along the integration); t beyond the
// integrated interval gives NaN.
template <class M>
__global__ void __launch_bounds__(256)
    k_dense_eval(const double* __restrict__ rec, long long cap,
                 const long long* __restrict__ nrec, long long n,
                 const double* __restrict__ t, int per_orbit, int T,
                 double* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * T) return;
  const long long i = idx / T;
  const int j = (int)(idx - i * T);
  const double tq = per_orbit ? t[i * T + j] : t[j];
  double* o = out + idx * 6;
  const long long nr = nrec[i];
  const double* r0 = rec + i * cap * kDenseRec;
  const double nanv = __longlong_as_double(0x7ff8000000000000ULL);
  if (nr <= 0) {
    for (int m = 0; m < 6; ++m) o[m] = nanv;
    return;
  }
  const double dir = r0[1] >= 0.0 ? 1.0 : -1.0;
  long long lo = 0, hi = nr - 1;
  while (lo < hi) {
    const long long mid = (lo + hi + 1) >> 1;
    if (dir * r0[mid * kDenseRec] <= dir * tq)
      lo = mid;
    else
      hi = mid - 1;
  }
  const double* r = r0 + lo * kDenseRec;
  const double h = r[1];
  double th = (tq - r[0]) / h;
  constexpr double kSlack = 1e-12;  // rounding of t at the interval's ends
  if (!(th >= -kSlack && th <= 1.0 + kSlack)) {
    for (int m = 0; m < 6; ++m) o[m] = nanv;
    return;
  }
  th = fmin(fmax(th, 0.0), 1.0);
  double wq[M::SF], wp[M::SF];
  dense_weights<M>(th, wq, wp);
  double sq[3] = {0.0, 0.0, 0.0}, sp[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < M::SF; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double gl = r[8 + 3 * l + m];
      sq[m] = fma(wq[l], gl, sq[m]);
      sp[m] = fma(wp[l], gl, sp[m]);
    }
  const double ht = h * th, hht = h * ht;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    o[m] = fma(hht, sq[m], fma(ht, r[5 + m], r[2 + m]));
    o[3 + m] = fma(ht, sp[m], r[5 + m]);
  }
}

// gx_integrate's argument checks and kernel arguments (GX_OK with n == 0:
// nothing to do).
int setup_args(const gx_pot* pot, const gx_solver_cfg* cfg, const double* w0,
               const double* t0, const double* t1, int32_t t_stride,
               const double* ts, int32_t T, int32_t ts_mode, int64_t n,
               const int64_t* order, double* out, int32_t* status,
               int64_t* nsteps, int64_t* naccept, void* workspace,
               IArgs& A) {
  if (!pot || !cfg || n < 0 || T < 1) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!w0 || !t0 || !t1 || !out || !workspace) return GX_EINVAL;
  if (t_stride != 0 && t_stride != 1) return GX_EINVAL;
  if (ts_mode < GX_TS_SHARED || ts_mode > GX_TS_FINAL) return GX_EINVAL;
  if ((ts_mode == GX_TS_SHARED || ts_mode == GX_TS_PER_ORBIT) && !ts)
    return GX_EINVAL;
  if (ts_mode == GX_TS_FINAL && T != 1) return GX_EINVAL;
  if (cfg->method != GX_DOPRI8 && cfg->method != GX_DOPRI5) return GX_EINVAL;
  if (cfg->controller != GX_CONSTANT && cfg->controller != GX_PID)
    return GX_EINVAL;
  if (cfg->controller == GX_CONSTANT && !(cfg->dt0 > 0.0)) return GX_EINVAL;
  if (cfg->controller == GX_PID &&
      !(cfg->rtol >= 0.0 && cfg->atol >= 0.0 && cfg->rtol + cfg->atol > 0.0 &&
        cfg->safety > 0.0 && cfg->factormin > 0.0 && cfg->factormax >= 1.0))
    return GX_EINVAL;

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
  A.root8 = (A.c1 == 0.125 && A.c2 == 0.0 && A.c3 == 0.0) ? 1 : 0;
  // "unbounded" still ends: every lane stops after kStepCap attempts
  A.max_steps = (cfg->max_steps > 0 && cfg->max_steps < kStepCap)
                    ? cfg->max_steps
                    : kStepCap;
  A.force_dtmin = cfg->force_dtmin;
  A.t_stride = t_stride;
  A.T = T;
  A.ts_mode = ts_mode;
  A.w0 = w0;
  A.t0 = t0;
  A.t1 = t1;
  A.ts = ts;
  A.n = n;
  A.order = reinterpret_cast<const long long*>(order);
  A.out = out;
  A.status = status;
  A.nsteps = reinterpret_cast<long long*>(nsteps);
  A.naccept = reinterpret_cast<long long*>(naccept);
  A.counter = static_cast<unsigned long long*>(workspace);
  A.h0 = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256);
  A.a0 = A.h0 + n;
  return GX_OK;
}

// The allocation-free entry point's workspace: a fixed array in device
// memory, one copy per GPU for the life of the process.  Host code never
// uses the array's name as an address: each device's copy is looked up once
// with hipGetSymbolAddress (with that device current) and that pointer is
// what kernels receive.  Nothing in it needs clearing between calls: every
// gx_integrate launch zeroes the queue counter itself (k_reset_queue) and
// the first-step prepass writes each orbit's h0 and a0 before the
// integrator reads them.
constexpr int64_t kScratchBytes = (int64_t)4 << 20;  // 4 MiB: 131,064 orbits
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return GX_ELAUNCH;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return GX_ELAUNCH;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int64_t gx_integrate_workspace_bytes(int64_t n) {
  return n < 0 ? 0 : 256 + 32 * n;
}

int gx_integrate(const gx_pot* pot, const gx_solver_cfg* cfg,
                 const double* w0, const double* t0, const double* t1,
                 int32_t t_stride, const double* ts, int32_t T,
                 int32_t ts_mode, int64_t n, const int64_t* order,
                 double* out, int32_t* status, int64_t* nsteps,
                 int64_t* naccept, void* workspace, void* stream) {
  IArgs A{};
  const int rc = setup_args(pot, cfg, w0, t0, t1, t_stride, ts, T, ts_mode, n,
                            order, out, status, nsteps, naccept, workspace, A);
  if (rc != GX_OK || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool pid = cfg->controller == GX_PID;
  using Full = EvalTable<true>;
  using Basic = EvalTable<false>;
  if (cfg->method == GX_DOPRI8) {
    if (pot->sig == kSigMilkyWay)
      return pid ? launch<Dopri8, true, EvalMilkyWay>(pot, A, s)
                 : launch<Dopri8, false, EvalMilkyWay>(pot, A, s);
    if (pot->sig == kSigMilkyWayRotating)
      return pid ? launch<Dopri8, true, EvalMilkyWayRotating>(pot, A, s)
                 : launch<Dopri8, false, EvalMilkyWayRotating>(pot, A, s);
    if (pot->sig == kSigMilkyWayTracked)
      return pid ? launch<Dopri8, true, EvalMilkyWayTracked>(pot, A, s)
                 : launch<Dopri8, false, EvalMilkyWayTracked>(pot, A, s);
    if (pot->sig == kSigMilkyWayPieces)
      return pid ? launch<Dopri8, true, EvalMilkyWayPieces>(pot, A, s)
                 : launch<Dopri8, false, EvalMilkyWayPieces>(pot, A, s);
    if (pid)
      return pot->basic ? launch<Dopri8, true, Basic>(pot, A, s)
                        : launch<Dopri8, true, Full>(pot, A, s);
    return pot->basic ? launch<Dopri8, false, Basic>(pot, A, s)
                      : launch<Dopri8, false, Full>(pot, A, s);
  }
  if (pid)
    return pot->basic ? launch<Dopri5, true, Basic>(pot, A, s)
                      : launch<Dopri5, true, Full>(pot, A, s);
  return pot->basic ? launch<Dopri5, false, Basic>(pot, A, s)
                    : launch<Dopri5, false, Full>(pot, A, s);
}

int gx_integrate_dense(const gx_pot* pot, const gx_solver_cfg* cfg,
                       const double* w0, const double* t0, const double* t1,
                       int32_t t_stride, const double* ts, int32_t T,
                       int32_t ts_mode, int64_t n, const int64_t* order,
                       double* out, int32_t* status, int64_t* nsteps,
                       int64_t* naccept, double* rec, int64_t rec_cap,
                       int64_t* nrec, void* workspace, void* stream) {
  if (!rec || !nrec || rec_cap < 1) return n == 0 ? GX_OK : GX_EINVAL;
  if (n > 0 && rec_cap > ((int64_t)1 << 40) / n) return GX_EINVAL;
  IArgs A{};
  hipStream_t s = (hipStream_t)stream;
  const int rc = setup_args(pot, cfg, w0, t0, t1, t_stride, ts, T, ts_mode, n,
                            order, out, status, nsteps, naccept, workspace, A);
  if (rc != GX_OK || n == 0) return rc;
  A.rec = rec;
  A.rec_cap = rec_cap;
  A.nrec = reinterpret_cast<long long*>(nrec);
  const bool pid = cfg->controller == GX_PID;
  using Full = EvalTable<true>;
  using Basic = EvalTable<false>;
  if (cfg->method == GX_DOPRI8) {
    if (pot->sig == kSigMilkyWay)
      return pid ? launch<Dopri8, true, EvalMilkyWay, true>(pot, A, s)
                 : launch<Dopri8, false, EvalMilkyWay, true>(pot, A, s);
    if (pid)
      return pot->basic ? launch<Dopri8, true, Basic, true>(pot, A, s)
                        : launch<Dopri8, true, Full, true>(pot, A, s);
    return pot->basic ? launch<Dopri8, false, Basic, true>(pot, A, s)
                      : launch<Dopri8, false, Full, true>(pot, A, s);
  }
  if (pid)
    return pot->basic ? launch<Dopri5, true, Basic, true>(pot, A, s)
                      : launch<Dopri5, true, Full, true>(pot, A, s);
  return pot->basic ? launch<Dopri5, false, Basic, true>(pot, A, s)
                    : launch<Dopri5, false, Full, true>(pot, A, s);
}

int gx_dense_eval(int32_t method, const double* rec, int64_t rec_cap,
                  const int64_t* nrec, int64_t n, const double* t,
                  int32_t per_orbit, int32_t T, double* out, void* stream) {
  if (n < 0 || T < 1 || rec_cap < 1) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!rec || !nrec || !t || !out) return GX_EINVAL;
  if (per_orbit != 0 && per_orbit != 1) return GX_EINVAL;
  if (n > ((int64_t)1 << 40) / T) return GX_EINVAL;
  const long long total = (long long)n * T;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return GX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const long long* nr = reinterpret_cast<const long long*>(nrec);
  if (method == GX_DOPRI8)
    hipLaunchKernelGGL(k_dense_eval<Dopri8>, dim3((unsigned)blocks), dim3(256),
                       0, s, rec, (long long)rec_cap, nr, (long long)n, t,
                       per_orbit, T, out);
  else if (method == GX_DOPRI5)
    hipLaunchKernelGGL(k_dense_eval<Dopri5>, dim3((unsigned)blocks), dim3(256),
                       0, s, rec, (long long)rec_cap, nr, (long long)n, t,
                       per_orbit, T, out);
  else
    return GX_EINVAL;
  return check(hipGetLastError());
}

int64_t gx_integrate_noalloc_max_orbits(void) {
  return (kScratchBytes - 256) / 32;
}

int gx_integrate_noalloc(const gx_pot* pot, const gx_solver_cfg* cfg,
                         const double* w0, const double* t0, const double* t1,
                         int32_t t_stride, const double* ts, int32_t T,
                         int32_t ts_mode, int64_t n, const int64_t* order,
                         double* out, int32_t* status, int64_t* nsteps,
                         int64_t* naccept, void* stream) {
  if (n > gx_integrate_noalloc_max_orbits()) return GX_ETOOBIG;
  if (n <= 0)  // nothing to launch; gx_integrate's argument checks
    return gx_integrate(pot, cfg, w0, t0, t1, t_stride, ts, T, ts_mode, n,
                        order, out, status, nsteps, naccept, nullptr, stream);
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != GX_OK) return rc;
  return gx_integrate(pot, cfg, w0, t0, t1, t_stride, ts, T, ts_mode, n, order,
                      out, status, nsteps, naccept, ws, stream);
}

}  // extern "C"
