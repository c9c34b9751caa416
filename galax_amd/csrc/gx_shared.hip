// gx_shared.hip -- one step sequence shared by a batch of orbits.
//
// galax's Integrator solves a batch whose start and end times are scalars
// as ONE ordinary differential equation over the (B, 6) state
// (legacy/integrator.py:269-355, orbit/solver.py:774-812): one step size,
// one accept/reject decision per step, the error norm the RMS over all 6B
// components, the automatic first step from norms over the whole batch.
// gd.evaluate_orbit takes this path for its second stage
// (legacy/funcs.py:42-213).  This file reproduces those semantics on the
// GPU; gx_integrate.hip gives every orbit its own steps instead.
//
// Design: one attempted step per launch, or per trip of one cooperative
// launch when the batch's workgroups all fit on the device at once.  Each lane evaluates the
// method for one orbit (the same code as the independent integrator,
// gx_rk.hpp); each workgroup reduces its error sum into a partial; the last
// workgroup to finish (a device-scope ticket) reduces the partials in a fixed
// order and runs the controller for the whole batch, writing a small control
// record that the next launch reads.  No lane ever waits on another
// workgroup, so a launch always drains.  The state (q, p and the step
// start's acceleration) is double-buffered: an attempted step writes its
// end state and end stage to the other buffer, and acceptance swaps the
// buffers.  Save times inside an attempted step are written from its dense
// output before the decision is known: a rejected attempt's values are
// overwritten by the accepted step that later covers those times, and an
// early stop marks every save past the last accepted step NaN.
// The host launches steps in batches and reads the record's `done` flag
// between batches; launches after the end return at once.
#include <hip/hip_cooperative_groups.h>

#include <cmath>

#include "gx_internal.hpp"
#include "gx_rk.hpp"

namespace gx {
namespace {

constexpr int kShBlock = 256;
constexpr int kShWaves = kShBlock / 64;
constexpr long long kShStepCap = 10000000;

// The control record (device memory, first 256 bytes of the workspace).
struct SCtl {
  double tau, tnext, tau1;  // solver time, end of the next attempt, end time
  double h0, d1;            // automatic first step: trial step, |f| norm
  double prev_inv, prev_prev_inv;
  long long nsteps, naccept;
  int status;
  int si;   // save times before si are final (at or before tau)
  int cur;  // buffer that holds the current state
  int at_dtmin, finished, done;
};
static_assert(sizeof(SCtl) <= 256, "control record");

struct SArgs {
  double dt0, rtol, atol, safety, factormin, factormax, dtmin, dtmax;
  double c1, c2, c3;  // PID exponents (i+p+d)/k, -(p+2d)/k, d/k
  long long max_steps;
  int force_dtmin;
  int T;
  double dir, tau0, tau1;
  const double* w0;
  const double* ts;
  long long n;
  double* out;
  int* status;
  long long* nsteps;
  long long* naccept;
  SCtl* ctl;
  unsigned int* counter;
  double* part;  // two partial sums per workgroup, in two sets (the
                 // cooperative solve alternates them by trip)
  double* y;     // two state buffers of n x 9: q, p (solver variables), and
                 // a(q, tau), the step start's first stage
  double* rec;   // dense records [n][rec_cap][kDenseRec], or null
  long long rec_cap;
  long long* nrec;
};

constexpr int kRow = 9;  // doubles per orbit in a state buffer

__device__ __forceinline__ void write_save(const SArgs& A, long long i, int s,
                                           const double (&q)[3],
                                           const double (&p)[3]) {
  double* o = A.out + (i * A.T + s) * 6;
  o[0] = q[0];
  o[1] = q[1];
  o[2] = q[2];
  o[3] = A.dir * p[0];
  o[4] = A.dir * p[1];
  o[5] = A.dir * p[2];
}

__device__ __forceinline__ void load_state(const double* y, long long i,
                                           double (&q)[3], double (&p)[3],
                                           double (&a)[3]) {
  const double* r = y + kRow * i;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    q[m] = r[m];
    p[m] = r[m + 3];
    a[m] = r[m + 6];
  }
}

__device__ __forceinline__ void store_state(double* y, long long i,
                                            const double (&q)[3],
                                            const double (&p)[3],
                                            const double (&a)[3]) {
  double* r = y + kRow * i;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    r[m] = q[m];
    r[m + 3] = p[m];
    r[m + 6] = a[m];
  }
}

// Sum of v over the workgroup, in a fixed order (every thread gets it).
__device__ __forceinline__ double block_sum(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kShWaves; ++w) s += lds[w];
  __syncthreads();
  return s;
}

// Publish this workgroup's two partial sums and return true in the last
// workgroup to do so; that workgroup then sums every partial, in index
// order, into (s0, s1).
__device__ bool last_block_sums(const SArgs& A, double a, double b,
                                double& s0, double& s1) {
  __shared__ double lds[kShWaves];
  __shared__ int is_last;
  a = block_sum(a, lds);
  b = block_sum(b, lds);
  if (threadIdx.x == 0) {
    A.part[2 * blockIdx.x] = a;
    A.part[2 * blockIdx.x + 1] = b;
    __threadfence();
    const unsigned int t = atomicAdd(A.counter, 1u);
    is_last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double x = 0.0, z = 0.0;
  for (unsigned int k = threadIdx.x; k < gridDim.x; k += kShBlock) {
    x += __hip_atomic_load(&A.part[2 * k], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    z += __hip_atomic_load(&A.part[2 * k + 1], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
  s0 = block_sum(x, lds);
  s1 = block_sum(z, lds);
  return true;
}

// Copy the initial states in (velocities times dir) with their start
// accelerations (the first step's first stage, by the step's own
// evaluator), write the saves at the start time, initialise the control
// record.  The first step is dt0 if given, else dtmin if the controller sets
// one (as the independent integrator), else the automatic rule's launches
// set it.
template <class E>
__global__ void __launch_bounds__(kShBlock) k_sh_setup(DPot P, SArgs A) {
  const long long i = (long long)blockIdx.x * kShBlock + threadIdx.x;
  int si = 0;
  while (si < A.T && A.dir * A.ts[si] <= A.tau0) ++si;
  const bool same = A.tau0 == A.tau1;
  if (i < A.n) {
    const double* w = A.w0 + 6 * i;
    const double q[3] = {w[0], w[1], w[2]};
    const double p[3] = {A.dir * w[3], A.dir * w[4], A.dir * w[5]};
    double a[3] = {0.0, 0.0, 0.0};
    if (!same) E::acc(P, q[0], q[1], q[2], A.dir * A.tau0, a[0], a[1], a[2]);
    store_state(A.y, i, q, p, a);
    const int ns = same ? A.T : si;
    for (int s = 0; s < ns; ++s) write_save(A, i, s, q, p);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SCtl c{};
    c.tau = A.tau0;
    c.tau1 = A.tau1;
    c.prev_inv = 1.0;
    c.prev_prev_inv = 1.0;
    c.status = GX_SUCCESS;
    c.si = same ? A.T : si;
    c.finished = same ? 1 : 0;
    c.done = same ? 1 : 0;
    if (A.dt0 > 0.0 || A.dtmin > 0.0) {
      double h = A.dt0 > 0.0 ? A.dt0 : A.dtmin;
      h = fmin(h, A.dtmax);
      if (A.dtmin > 0.0) {
        c.at_dtmin = h <= A.dtmin ? 1 : 0;
        h = fmax(h, A.dtmin);
      }
      c.tnext = fmin(A.tau0 + h, A.tau1);
    }
    *A.ctl = c;
    *A.counter = 0u;
  }
}

// diffrax's automatic first step over the whole batch (Hairer, Norsett &
// Wanner II.4), in two launches: the norms of y and f(y) and a trial step,
// then the norm of the change of f over the trial step.
__global__ void __launch_bounds__(kShBlock) k_sh_first_a(SArgs A) {
  const long long i = (long long)blockIdx.x * kShBlock + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  if (i < A.n) {
    double q[3], p[3], a[3];
    load_state(A.y, i, q, p, a);
    for (int m = 0; m < 3; ++m) {
      const double sq = A.atol + fabs(q[m]) * A.rtol;
      const double sp = A.atol + fabs(p[m]) * A.rtol;
      s0 += (q[m] / sq) * (q[m] / sq) + (p[m] / sp) * (p[m] / sp);
      s1 += (p[m] / sq) * (p[m] / sq) + (a[m] / sp) * (a[m] / sp);
    }
  }
  double t0, t1;
  if (!last_block_sums(A, s0, s1, t0, t1)) return;
  if (threadIdx.x == 0) {
    const double n6 = 6.0 * (double)A.n;
    const double d0 = sqrt(t0 / n6), d1 = sqrt(t1 / n6);
    const bool cond = (d0 < 1e-5) || (d1 < 1e-5);
    A.ctl->h0 = cond ? 1e-6 : 0.01 * (d0 / d1);
    A.ctl->d1 = d1;
    *A.counter = 0u;
  }
}

template <class E>
__global__ void __launch_bounds__(kShBlock) k_sh_first_b(DPot P, SArgs A,
                                                         double order) {
  const long long i = (long long)blockIdx.x * kShBlock + threadIdx.x;
  const double h0 = A.ctl->h0;
  double s2 = 0.0;
  if (i < A.n) {
    double q[3], p[3], a0[3], a1[3], q1[3], p1[3];
    load_state(A.y, i, q, p, a0);
    for (int m = 0; m < 3; ++m) {
      q1[m] = q[m] + p[m] * h0;
      p1[m] = p[m] + a0[m] * h0;
    }
    E::acc(P, q1[0], q1[1], q1[2], A.dir * (A.tau0 + h0), a1[0], a1[1],
           a1[2]);
    for (int m = 0; m < 3; ++m) {
      const double sq = A.atol + fabs(q[m]) * A.rtol;
      const double sp = A.atol + fabs(p[m]) * A.rtol;
      const double u = (p1[m] - p[m]) / sq;  // change of dq/dt = p
      const double v = (a1[m] - a0[m]) / sp;
      s2 += u * u + v * v;
    }
  }
  double t2, unused;
  if (!last_block_sums(A, s2, 0.0, t2, unused)) return;
  if (threadIdx.x == 0) {
    SCtl& c = *A.ctl;
    const double d2 = sqrt(t2 / (6.0 * (double)A.n)) / h0;
    const double maxd = fmax(c.d1, d2);
    const double h1 = (maxd <= 1e-15) ? fmax(1e-6, h0 * 1e-3)
                                      : pow(0.01 / maxd, 1.0 / (order + 1.0));
    double h = fmin(100.0 * h0, h1);
    h = fmin(h, A.dtmax);
    if (A.dtmin > 0.0) {
      c.at_dtmin = h <= A.dtmin ? 1 : 0;
      h = fmax(h, A.dtmin);
    }
    c.tnext = fmin(A.tau0 + h, A.tau1);
    *A.counter = 0u;
  }
}

__device__ __forceinline__ double clip_end(double tprev, double tnext,
                                           double t1, bool keep) {
  return (tnext > t1 - 1e-10) ? (keep ? t1 : tprev + 0.5 * (t1 - tprev))
                              : tnext;
}

// One attempted step of the whole batch for this lane's orbit, with the
// save times inside it written from the step's dense output (before the
// decision; see above): the lane's squared scaled error and a non-finite
// flag for the batch's norm.
template <class M, bool PID, class E>
__device__ __forceinline__ void sh_attempt(const DPot& P, const SArgs& A,
                                           const SCtl& c, double& sq,
                                           double& bad) {
  const long long i = (long long)blockIdx.x * kShBlock + threadIdx.x;
  const bool valid = i < A.n;
  const double dir = A.dir;
  const double* ycur = A.y + (long long)c.cur * kRow * A.n;
  double* yoth = A.y + (long long)(c.cur ^ 1) * kRow * A.n;
  sq = 0.0;
  bad = 0.0;
  if (valid && !c.finished) {
    constexpr int SF = M::SF;
    const double h = c.tnext - c.tau;
    double q[3], p[3], g[SF][3], q1[3], p1[3], eq[3], ep[3];
    load_state(ycur, i, q, p, g[0]);
    rk_step<M, E, PID>(P, dir, c.tau, h, c.tnext, q, p, g, q1, p1, eq, ep);
    rk_end_stage<M, E>(P, dir, c.tnext, q1, g);
    store_state(yoth, i, q1, p1, g[SF - 1]);
    // this attempt's record goes to the next free slot; a rejection leaves
    // the slot to the next attempt
    if (A.rec && c.naccept < A.rec_cap)
      write_record<M>(A.rec + (i * A.rec_cap + c.naccept) * kDenseRec, dir,
                      c.tau, h, q, p, g);
    bool finite = true;
    for (int m = 0; m < 3; ++m)
      finite = finite && isfinite(q1[m]) && isfinite(p1[m]);
    bad = finite ? 0.0 : 1.0;
    for (int s = c.si; s < A.T; ++s) {
      const double ts_tau = dir * A.ts[s];
      if (ts_tau > c.tnext) break;
      if (ts_tau == c.tnext) {
        write_save(A, i, s, q1, p1);
      } else {
        double qs[3], ps[3];
        dense_eval<M>(div(ts_tau - c.tau, h), h, q, p, g, qs, ps);
        write_save(A, i, s, qs, ps);
      }
    }
    if (PID) {
      for (int m = 0; m < 3; ++m) {
        const double uq = eq[m] / (A.atol + fmax(fabs(q[m]), fabs(q1[m])) * A.rtol);
        const double up = ep[m] / (A.atol + fmax(fabs(p[m]), fabs(p1[m])) * A.rtol);
        sq += uq * uq + up * up;
      }
    }
  }
}

// The controller for the whole batch: the record after the attempt that
// started from `c`, given the batch's summed squared scaled error SQ and
// non-finite count NB.
template <bool PID>
__device__ SCtl sh_control(const SArgs& A, const SCtl& c, double SQ, double NB) {
  SCtl n = c;
  if (c.finished) {
    n.done = 1;
    return n;
  }
  const double dir = A.dir;
  n.nsteps = c.nsteps + 1;
  const double h = c.tnext - c.tau;
  if (NB > 0.0 || !(h >= 0.0)) {
    n.status = GX_NONFINITE;
    n.finished = 1;
    return n;
  }
  bool keep = true;
  double next_t1;
  if (PID) {
    const double scaled = sqrt(SQ / (6.0 * (double)A.n));
    keep = scaled < 1.0;
    if (A.dtmin > 0.0) keep = keep || c.at_dtmin;
    const double inv = 1.0 / scaled;
    const double f1 = (A.c1 == 0.0) ? 1.0 : pow(inv, A.c1);
    const double f2 = (A.c2 == 0.0) ? 1.0 : pow(c.prev_inv, A.c2);
    const double f3 = (A.c3 == 0.0) ? 1.0 : pow(c.prev_prev_inv, A.c3);
    const double fmin_ = keep ? 1.0 : A.factormin;
    const double factor = fmin(fmax(A.safety * f1 * f2 * f3, fmin_), A.factormax);
    double dt = fmin(h * factor, A.dtmax);
    if (A.dtmin > 0.0) {
      if (!A.force_dtmin && dt < A.dtmin) n.status = GX_DT_MIN_REACHED;
      n.at_dtmin = dt <= A.dtmin ? 1 : 0;
      dt = fmax(dt, A.dtmin);
    }
    if (keep) {
      n.prev_prev_inv = c.prev_inv;
      n.prev_inv = (inv == 0.0 || isinf(inv)) ? 1.0 : inv;
    }
    next_t1 = (keep ? c.tnext : c.tau) + dt;
  } else {
    next_t1 = c.tnext + A.dt0;
  }
  if (keep) {
    n.cur = c.cur ^ 1;
    n.naccept = c.naccept + 1;
    if (A.rec && n.naccept > A.rec_cap) {
      n.status = GX_MAX_STEPS_REACHED;  // out of record slots
      n.finished = 1;
    }
    int s = c.si;
    while (s < A.T && dir * A.ts[s] <= c.tnext) ++s;
    n.si = s;
  }
  const double tprev = fmin(keep ? c.tnext : c.tau, c.tau1);
  n.tnext = clip_end(tprev, next_t1, c.tau1, keep);
  n.tau = tprev;
  if (n.status == GX_DT_MIN_REACHED || n.tau >= c.tau1)
    n.finished = 1;
  else if (n.nsteps >= A.max_steps) {
    n.status = GX_MAX_STEPS_REACHED;
    n.finished = 1;
  }
  return n;
}

// One attempted step of the whole batch per launch: the last workgroup to
// publish its partial sums runs the controller and writes the record.
template <class M, bool PID, class E>
__global__ void __launch_bounds__(kShBlock) k_sh_step(DPot P, SArgs A) {
  const SCtl c = *A.ctl;
  if (c.done) return;
  double sq, bad;
  sh_attempt<M, PID, E>(P, A, c, sq, bad);
  double SQ, NB;
  if (!last_block_sums(A, sq, bad, SQ, NB)) return;
  if (threadIdx.x != 0) return;
  *A.ctl = sh_control<PID>(A, c, SQ, NB);
  *A.counter = 0u;
}

// The whole solve in one cooperative launch (every workgroup resident):
// each trip is one attempted step; the workgroups publish their partial
// sums (double-buffered by trip parity, so that a workgroup already on the
// next trip cannot overwrite sums another is still reading), meet at one
// grid barrier, and then every workgroup sums the partials in index order
// -- the same order and arithmetic as last_block_sums -- and runs the
// controller itself, so all take the same decision without a second
// barrier.  Workgroup 0 also writes the record for the host and the finish
// kernel.  Every workgroup leaves on the same trip (the record's done).
template <class M, bool PID, class E>
__global__ void __launch_bounds__(kShBlock) k_sh_persist(DPot P, SArgs A) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  __shared__ SCtl cs;
  __shared__ double lds[kShWaves];
  if (threadIdx.x == 0) cs = *A.ctl;
  __syncthreads();
  for (long long trip = 0;; ++trip) {
    const SCtl c = cs;
    if (c.done) break;
    double sq, bad;
    sh_attempt<M, PID, E>(P, A, c, sq, bad);
    const double a = block_sum(sq, lds), b = block_sum(bad, lds);
    double* part = A.part + 2 * (long long)gridDim.x * (trip & 1);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = a;
      part[2 * blockIdx.x + 1] = b;
      __threadfence();
    }
    grid.sync();
    double x = 0.0, z = 0.0;
    for (unsigned int k = threadIdx.x; k < gridDim.x; k += kShBlock) {
      x += __hip_atomic_load(&part[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      z += __hip_atomic_load(&part[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const double SQ = block_sum(x, lds), NB = block_sum(z, lds);
    if (threadIdx.x == 0) {
      const SCtl n = sh_control<PID>(A, c, SQ, NB);
      cs = n;
      if (blockIdx.x == 0) *A.ctl = n;
    }
    __syncthreads();
  }
}

// Per-orbit results: NaN for save times never reached; the batch's status
// and step counts for every orbit.
__global__ void __launch_bounds__(kShBlock) k_sh_finish(SArgs A) {
  const long long i = (long long)blockIdx.x * kShBlock + threadIdx.x;
  if (i >= A.n) return;
  const SCtl c = *A.ctl;
  const double nanv = __longlong_as_double(0x7ff8000000000000ULL);
  const double nq[3] = {nanv, nanv, nanv};
  for (int s = c.si; s < A.T; ++s) {
    double* o = A.out + (i * A.T + s) * 6;
    for (int m = 0; m < 6; ++m) o[m] = nq[0];
  }
  if (A.status) A.status[i] = c.status;
  if (A.nsteps) A.nsteps[i] = c.nsteps;
  if (A.naccept) A.naccept[i] = c.naccept;
  if (A.nrec) A.nrec[i] = c.naccept < A.rec_cap ? c.naccept : A.rec_cap;
}

// Two ints of pinned host memory per host thread for the done flags
// (allocated on first use, kept until the process ends).
int* pinned_flags() {
  thread_local int* flags = nullptr;
  if (!flags) {
    void* p = nullptr;
    if (hipHostMalloc(&p, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess)
      return nullptr;
    flags = static_cast<int*>(p);
  }
  return flags;
}

// A stream of this host thread used only to capture the step graph (the
// graph is then launched on the caller's stream); default priority.
hipStream_t capture_stream() {
  thread_local hipStream_t s = nullptr;
  if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    s = nullptr;
  return s;
}

constexpr int kGraphSteps = 32;

template <class M, bool PID, class E>
int run(const gx_pot* pot, SArgs& A, bool first_step, hipStream_t stream) {
  const unsigned blocks = (unsigned)((A.n + kShBlock - 1) / kShBlock);
  hipLaunchKernelGGL(k_sh_setup<E>, dim3(blocks), dim3(kShBlock), 0, stream,
                     pot->dev, A);
  if (first_step) {
    hipLaunchKernelGGL(k_sh_first_a, dim3(blocks), dim3(kShBlock), 0, stream,
                       A);
    hipLaunchKernelGGL(k_sh_first_b<E>, dim3(blocks), dim3(kShBlock), 0,
                       stream, pot->dev, A, (double)M::order);
  }
  if (hipGetLastError() != hipSuccess) return GX_ELAUNCH;
  // The whole solve in one cooperative launch when every workgroup of the
  // batch can be resident at once (the launch is refused otherwise, and the
  // step launches below take over).  The call stays synchronous.
  {
    const auto pk = k_sh_persist<M, PID, E>;
    int coop = 0, occ = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch,
                              pot->device) == hipSuccess &&
        coop &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, pk, kShBlock, 0) ==
            hipSuccess &&
        (long long)blocks <= (long long)occ * num_cus(pot->device)) {
      DPot Pd = pot->dev;
      SArgs Ad = A;
      void* args[] = {&Pd, &Ad};
      if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(pk),
                                     dim3(blocks), dim3(kShBlock), args, 0,
                                     stream) == hipSuccess) {
        hipLaunchKernelGGL(k_sh_finish, dim3(blocks), dim3(kShBlock), 0,
                           stream, A);
        if (hipGetLastError() != hipSuccess) return GX_ELAUNCH;
        return hipStreamSynchronize(stream) == hipSuccess ? GX_OK : GX_ELAUNCH;
      }
      (void)hipGetLastError();
    }
  }
  // Step launches go out in batches; after each batch the record's `done`
  // flag is copied to pinned host memory behind an event.  Two batches are
  // in flight, so the GPU never waits for the host's check.  Every launch
  // that does not find the record done attempts a step, and the record is
  // done after at most max_steps + 2 attempts.  The first batch is launched
  // directly (short solves end there); later batches replay one captured
  // HIP graph of kGraphSteps step launches, which costs the host a fraction
  // of a launch per step.
  int* flags = pinned_flags();
  if (!flags) return GX_ENOMEM;
  hipGraphExec_t gexec = nullptr;
  {
    hipStream_t cap = capture_stream();
    hipGraph_t graph = nullptr;
    bool ok = cap && hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) ==
                         hipSuccess;
    if (ok) {
      for (int k = 0; k < kGraphSteps; ++k)
        hipLaunchKernelGGL((k_sh_step<M, PID, E>), dim3(blocks),
                           dim3(kShBlock), 0, cap, pot->dev, A);
      ok = hipStreamEndCapture(cap, &graph) == hipSuccess && graph;
    }
    if (ok && hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) !=
                  hipSuccess)
      gexec = nullptr;
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();  // a failed capture leaves the direct launches
  }
  hipEvent_t ev[2];
  if (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess)
    return GX_ELAUNCH;
  if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return GX_ELAUNCH;
  }
  const long long limit = A.max_steps + 2;
  long long launched = 0;
  int chunk = 8;
  int rc = GX_OK;
  auto issue = [&](int slot) {
    flags[slot] = 0;
    if (gexec && launched > 0) {
      if (hipGraphLaunch(gexec, stream) != hipSuccess) rc = GX_ELAUNCH;
      launched += kGraphSteps;
    } else {
      for (int k = 0; k < chunk; ++k)
        hipLaunchKernelGGL((k_sh_step<M, PID, E>), dim3(blocks),
                           dim3(kShBlock), 0, stream, pot->dev, A);
      launched += chunk;
      chunk = chunk < 64 ? 2 * chunk : 64;
    }
    if (hipMemcpyAsync(&flags[slot], &A.ctl->done, sizeof(int),
                       hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipEventRecord(ev[slot], stream) != hipSuccess)
      rc = GX_ELAUNCH;
  };
  int slot = 0;
  issue(0);
  while (rc == GX_OK) {
    issue(slot ^ 1);
    if (hipEventSynchronize(ev[slot]) != hipSuccess) {
      rc = GX_ELAUNCH;
      break;
    }
    if (flags[slot]) break;
    if (launched > limit + 128) rc = GX_ELAUNCH;  // the record never ended
    slot ^= 1;
  }
  // the other batch's copy targets the same pinned flags: let it land
  (void)hipEventSynchronize(ev[slot ^ 1]);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (rc != GX_OK) return rc;
  hipLaunchKernelGGL(k_sh_finish, dim3(blocks), dim3(kShBlock), 0, stream, A);
  return check(hipGetLastError());
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int64_t gx_integrate_shared_workspace_bytes(int64_t n) {
  if (n < 0) return 0;
  const int64_t blocks = (n + kShBlock - 1) / kShBlock;
  return 512 + 32 * blocks + 8 * (2 * kRow * n);
}

}  // extern "C"

namespace {

int shared_solve(const gx_pot* pot, const gx_solver_cfg* cfg,
                 const double* w0, double t0, double t1, const double* ts,
                 int32_t T, int64_t n, double* out, int32_t* status,
                 int64_t* nsteps, int64_t* naccept, double* rec,
                 int64_t rec_cap, int64_t* nrec, void* workspace,
                 void* stream) {
  if (!pot || !cfg || n < 0 || T < 1) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!w0 || !ts || !out || !workspace) return GX_EINVAL;
  if (!std::isfinite(t0) || !std::isfinite(t1)) return GX_EINVAL;
  if (n > ((int64_t)1 << 40)) return GX_EINVAL;
  if (cfg->method != GX_DOPRI8 && cfg->method != GX_DOPRI5) return GX_EINVAL;
  if (cfg->controller != GX_CONSTANT && cfg->controller != GX_PID)
    return GX_EINVAL;
  if (cfg->controller == GX_CONSTANT && !(cfg->dt0 > 0.0)) return GX_EINVAL;
  if (cfg->controller == GX_PID &&
      !(cfg->rtol >= 0.0 && cfg->atol >= 0.0 && cfg->rtol + cfg->atol > 0.0 &&
        cfg->safety > 0.0 && cfg->factormin > 0.0 && cfg->factormax >= 1.0))
    return GX_EINVAL;
  const bool pid = cfg->controller == GX_PID;
  SArgs A{};
  A.dt0 = cfg->dt0;
  A.rtol = cfg->rtol;
  A.atol = cfg->atol;
  A.safety = cfg->safety;
  A.factormin = cfg->factormin;
  A.factormax = cfg->factormax;
  A.dtmin = (pid && cfg->dtmin > 0.0) ? cfg->dtmin : 0.0;
  A.dtmax = (pid && cfg->dtmax > 0.0) ? cfg->dtmax : INFINITY;
  const double korder = cfg->method == GX_DOPRI8 ? 8.0 : 5.0;
  A.c1 = (cfg->icoeff + cfg->pcoeff + cfg->dcoeff) / korder;
  A.c2 = -(cfg->pcoeff + 2.0 * cfg->dcoeff) / korder;
  A.c3 = cfg->dcoeff / korder;
  A.max_steps = (cfg->max_steps > 0 && cfg->max_steps < kShStepCap)
                    ? cfg->max_steps
                    : kShStepCap;
  A.force_dtmin = cfg->force_dtmin;
  A.T = T;
  A.dir = (t0 <= t1) ? 1.0 : -1.0;
  A.tau0 = A.dir * t0;
  A.tau1 = A.dir * t1;
  A.w0 = w0;
  A.ts = ts;
  A.n = n;
  A.out = out;
  A.status = status;
  A.nsteps = reinterpret_cast<long long*>(nsteps);
  A.naccept = reinterpret_cast<long long*>(naccept);
  char* ws = static_cast<char*>(workspace);
  const int64_t blocks = (n + kShBlock - 1) / kShBlock;
  A.ctl = reinterpret_cast<SCtl*>(ws);
  A.counter = reinterpret_cast<unsigned int*>(ws + 256);
  A.part = reinterpret_cast<double*>(ws + 512);  // two sets of partials
  A.y = A.part + 4 * blocks;
  A.rec = rec;
  A.rec_cap = rec_cap;
  A.nrec = reinterpret_cast<long long*>(nrec);
  const bool first_step = !(cfg->dt0 > 0.0) && !(A.dtmin > 0.0) && t0 != t1;
  if (!(cfg->dt0 > 0.0) && !pid) return GX_EINVAL;

  hipStream_t s = (hipStream_t)stream;
  using Full = EvalTable<true>;
  using Basic = EvalTable<false>;
  if (cfg->method == GX_DOPRI8) {
    if (pot->sig == kSigMilkyWay)
      return pid ? run<Dopri8, true, EvalMilkyWay>(pot, A, first_step, s)
                 : run<Dopri8, false, EvalMilkyWay>(pot, A, first_step, s);
    if (pid)
      return pot->basic ? run<Dopri8, true, Basic>(pot, A, first_step, s)
                        : run<Dopri8, true, Full>(pot, A, first_step, s);
    return pot->basic ? run<Dopri8, false, Basic>(pot, A, first_step, s)
                      : run<Dopri8, false, Full>(pot, A, first_step, s);
  }
  if (pid)
    return pot->basic ? run<Dopri5, true, Basic>(pot, A, first_step, s)
                      : run<Dopri5, true, Full>(pot, A, first_step, s);
  return pot->basic ? run<Dopri5, false, Basic>(pot, A, first_step, s)
                    : run<Dopri5, false, Full>(pot, A, first_step, s);
}

}  // namespace

extern "C" {

int gx_integrate_shared(const gx_pot* pot, const gx_solver_cfg* cfg,
                        const double* w0, double t0, double t1,
                        const double* ts, int32_t T, int64_t n, double* out,
                        int32_t* status, int64_t* nsteps, int64_t* naccept,
                        void* workspace, void* stream) {
  return shared_solve(pot, cfg, w0, t0, t1, ts, T, n, out, status, nsteps,
                      naccept, nullptr, 0, nullptr, workspace, stream);
}

int gx_integrate_shared_dense(const gx_pot* pot, const gx_solver_cfg* cfg,
                              const double* w0, double t0, double t1,
                              const double* ts, int32_t T, int64_t n,
                              double* out, int32_t* status, int64_t* nsteps,
                              int64_t* naccept, double* rec, int64_t rec_cap,
                              int64_t* nrec, void* workspace, void* stream) {
  if (!rec || !nrec || rec_cap < 1) return n == 0 ? GX_OK : GX_EINVAL;
  if (n > 0 && rec_cap > ((int64_t)1 << 40) / n) return GX_EINVAL;
  return shared_solve(pot, cfg, w0, t0, t1, ts, T, n, out, status, nsteps,
                      naccept, rec, rec_cap, nrec, workspace, stream);
}

}  // extern "C"
