// gx_rk.hpp -- one step of an explicit Runge-Kutta-Nystrom method over one
// orbit, and the step's dense output, shared by the independent-orbit and
// the shared-step integrators.
#pragma once

#include <type_traits>

#include "gx_potential.hpp"
#include "gx_tableau.hpp"

namespace gx {

// Doubles per dense-output record of an accepted step: t, h, q, v and up to
// 14 stage accelerations (the ABI's GX_DENSE_RECORD).
constexpr int kDenseRec = 8 + 3 * 14;

// Evaluators whose field rides a track and that look the track up at all of
// a step's stage times when the step starts (E::kTrackPrefetch, with
// E::track_n and E::acc_at).
template <class E, class = void>
struct track_prefetch : std::false_type {};
template <class E>
struct track_prefetch<E, std::void_t<decltype(E::kTrackPrefetch)>>
    : std::bool_constant<E::kTrackPrefetch> {};

// The first stage j <= i whose node c_j equals stage i's (i itself when its
// node is new).  Stage times tb + c_i h of equal nodes are the same double,
// so their track positions are looked up once; the end stage (SF-1) sits at
// te, which is computed apart from the nodes, and is always its own.
template <class M>
constexpr int same_node(int i) {
  if (i == M::SF - 1) return i;
  for (int j = 1; j < i; ++j)
    if (M::kf.c[j] == M::kf.c[i]) return j;
  return i;
}

// Stages i0 + 1 .. i0 + N as one batch of lookups: the U stages with a new
// node (uniq), and for each stage its slot among them, or -1 - (j - 1) to
// copy the position of the earlier stage j with the same node.
template <class M, int i0, int N>
struct NodeBatch {
  struct Map {
    int uniq[N];
    int slot[N];
    int U;
  };
  static constexpr Map make() {
    Map m{};
    for (int j = 0; j < N; ++j) {
      const int i = i0 + 1 + j, s = same_node<M>(i);
      if (s == i) {
        m.uniq[m.U] = i;
        m.slot[j] = m.U++;
      } else {
        m.slot[j] = -1 - (s - 1);
      }
    }
    return m;
  }
  static constexpr Map map = make();
  static constexpr int U = map.U;
};

// The first NA stages (1 .. NA) hold half of the step's distinct stage times
// (rounded up): the split of the lookups into two batches of loads.
template <class M>
constexpr int half_split() {
  int total = 0;
  for (int i = 1; i < M::SF; ++i) total += same_node<M>(i) == i;
  int seen = 0;
  for (int i = 1; i < M::SF; ++i) {
    seen += same_node<M>(i) == i;
    if (2 * seen >= total) return i;
  }
  return M::SF - 1;
}

// Track positions at stage times i0 + 1 .. i0 + N of a step (the end stage
// at te), into d[i0 .. i0 + N - 1]; a stage whose node repeats an earlier
// stage's copies that stage's position (which an earlier batch, or this
// one, has filled).
template <class M, class E, int i0, int N>
__device__ __forceinline__ void stage_tracks(const DPot& P, double dir,
                                             double tb, double h, double te,
                                             double (&d)[M::SF - 1][3]) {
  using B = NodeBatch<M, i0, N>;
  constexpr int U = B::U;
  double tt[U], dd[U][3];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = B::map.uniq[u];
    tt[u] = dir * ((i == M::SF - 1) ? te : fma(M::kf.c[i], h, tb));
  }
  E::template track_n<U>(P, tt, dd);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int sl = B::map.slot[j];
#pragma unroll
    for (int m = 0; m < 3; ++m)
      d[i0 + j][m] = sl >= 0 ? dd[sl >= 0 ? sl : 0][m] : d[sl >= 0 ? 0 : -1 - sl][m];
  }
}

// One step of method M from (tb, q, p) with step h to te = tb + h (solver
// time tau; physical time dir * tau, velocities pre-multiplied by dir).
//
// g[0] holds a(q, tb) on entry (the previous step's end stage, or the
// orbit's start acceleration); the step fills the method's stages
// g[1..S-1].  Dopri5's last stage is its end stage a(q1, te) (first same as
// last); Dopri8's end stage (slot SF-1) is evaluated separately by
// rk_end_stage once the step is accepted, for the dense output and as the
// next step's first stage.  Either way q1 is the end stage's position: the
// FSAL row of the tableau is b, so the stage formula of that row is the
// solution formula in the same order of operations.
//
// Small evaluators unroll the stages (every index static).  Large ones keep
// the stage loop rolled so that the potential evaluation is emitted once;
// each stage's coefficients and its slot in g[][] are selected by a chain
// of compile-time-indexed branches on the wave-uniform stage number, which
// keeps g[][] in registers.
template <class M, class E, bool ERR>
__device__ __forceinline__ void rk_step(const DPot& P, double dir, double tb,
                                        double h, double te,
                                        const double (&q)[3],
                                        const double (&p)[3],
                                        double (&g)[M::SF][3],
                                        double (&q1)[3], double (&p1)[3],
                                        double (&eq)[3], double (&ep)[3]) {
  constexpr int S = M::S;
  constexpr int SF = M::SF;
  // Stage positions as q + rs (h p) + (h h) s: two FMAs per component
  const double hh = h * h;
  const double hp[3] = {h * p[0], h * p[1], h * p[2]};
  // track positions at the stage times 1 .. SF-1 (slot i-1 for stage i),
  // looked up in two batches of loads
  constexpr bool kPre = track_prefetch<E>::value;
  static_assert(!(kPre && E::kLarge), "track prefetch needs unrolled stages");
  double trk[kPre ? SF - 1 : 1][3];
  if constexpr (kPre) {
    constexpr int NA = half_split<M>();
    stage_tracks<M, E, 0, NA>(P, dir, tb, h, te, trk);
    stage_tracks<M, E, NA, SF - 1 - NA>(P, dir, tb, h, te, trk);
  }
  if constexpr (!E::kLarge) {
#pragma unroll
    for (int i = 1; i < S; ++i) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      bool any = false;
#pragma unroll
      for (int l = 0; l < i; ++l) {
        if (M::kf.aa[i][l] != 0.0) {
          const double a = M::kf.aa[i][l];
          s0 = fma(a, g[l][0], s0);
          s1 = fma(a, g[l][1], s1);
          s2 = fma(a, g[l][2], s2);
          any = true;
        }
      }
      const double rs = M::kf.rs[i];
      double Q0 = fma(rs, hp[0], q[0]), Q1 = fma(rs, hp[1], q[1]),
             Q2 = fma(rs, hp[2], q[2]);
      if (any) {
        Q0 = fma(hh, s0, Q0);
        Q1 = fma(hh, s1, Q1);
        Q2 = fma(hh, s2, Q2);
      }
      const double ts = (i == SF - 1) ? te : fma(M::kf.c[i], h, tb);
      if constexpr (kPre)
        E::acc_at(P, Q0, Q1, Q2, dir * ts, trk[i - 1][0], trk[i - 1][1],
                  trk[i - 1][2], g[i][0], g[i][1], g[i][2]);
      else
        E::acc(P, Q0, Q1, Q2, dir * ts, g[i][0], g[i][1], g[i][2]);
    }
    if constexpr (kPre && SF > S) {
      // the end stage's track position waits in its acceleration slot
      // until rk_end_stage evaluates it
      g[SF - 1][0] = trk[SF - 2][0];
      g[SF - 1][1] = trk[SF - 2][1];
      g[SF - 1][2] = trk[SF - 2][2];
    }
  } else {
    double Qx = q[0], Qy = q[1], Qz = q[2];
#pragma nounroll
    for (int i = 1; i < S; ++i) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, rs = 0.0, c = 0.0;
#pragma unroll
      for (int k = 1; k < S; ++k) {
        if (k == i) {
          asm volatile("");  // no if-conversion into selects
#pragma unroll
          for (int l = 0; l < k; ++l) {
            const double a = M::kf.aa[k][l];
            if (a != 0.0) {
              s0 = fma(a, g[l][0], s0);
              s1 = fma(a, g[l][1], s1);
              s2 = fma(a, g[l][2], s2);
            }
          }
          rs = M::kf.rs[k];
          c = M::kf.c[k];
        }
      }
      Qx = fma(hh, s0, fma(rs, hp[0], q[0]));
      Qy = fma(hh, s1, fma(rs, hp[1], q[1]));
      Qz = fma(hh, s2, fma(rs, hp[2], q[2]));
      const double ts = (i == SF - 1) ? te : fma(c, h, tb);
      double ax, ay, az;
      E::acc(P, Qx, Qy, Qz, dir * ts, ax, ay, az);
#pragma unroll
      for (int k = 1; k < S; ++k) {
        if (k == i) {
          asm volatile("");
          g[k][0] = ax;
          g[k][1] = ay;
          g[k][2] = az;
        }
      }
    }
  }
  // q1 by the end stage's row (= b), p1 and the error estimate
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0, ea0 = 0.0, ea1 = 0.0, ea2 = 0.0;
#pragma unroll
  for (int l = 0; l < SF - 1; ++l) {
    if (M::kf.aa[SF - 1][l] != 0.0) {
      const double a = M::kf.aa[SF - 1][l];
      a0 = fma(a, g[l][0], a0);
      a1 = fma(a, g[l][1], a1);
      a2 = fma(a, g[l][2], a2);
    }
  }
#pragma unroll
  for (int l = 0; l < S; ++l) {
    if (M::kf.b[l] != 0.0) {
      const double b = M::kf.b[l];
      b0 = fma(b, g[l][0], b0);
      b1 = fma(b, g[l][1], b1);
      b2 = fma(b, g[l][2], b2);
    }
    if (ERR) {
      if (M::kf.e[l] != 0.0) {
        const double e = M::kf.e[l];
        e0 = fma(e, g[l][0], e0);
        e1 = fma(e, g[l][1], e1);
        e2 = fma(e, g[l][2], e2);
      }
      if (M::kf.ea[l] != 0.0) {
        const double ea = M::kf.ea[l];
        ea0 = fma(ea, g[l][0], ea0);
        ea1 = fma(ea, g[l][1], ea1);
        ea2 = fma(ea, g[l][2], ea2);
      }
    }
  }
  constexpr double rse = M::kf.rs[SF - 1], se = M::kf.se;
  q1[0] = fma(hh, a0, fma(rse, hp[0], q[0]));
  q1[1] = fma(hh, a1, fma(rse, hp[1], q[1]));
  q1[2] = fma(hh, a2, fma(rse, hp[2], q[2]));
  p1[0] = fma(h, b0, p[0]);
  p1[1] = fma(h, b1, p[1]);
  p1[2] = fma(h, b2, p[2]);
  if (ERR) {
    ep[0] = h * e0;
    ep[1] = h * e1;
    ep[2] = h * e2;
    eq[0] = fma(hh, ea0, se * hp[0]);
    eq[1] = fma(hh, ea1, se * hp[1]);
    eq[2] = fma(hh, ea2, se * hp[2]);
  }
}

// The end stage a(q1, te) into slot SF-1 when it is not one of the
// method's own stages (Dopri8).  For a track-prefetching evaluator the slot
// holds the track position at te that rk_step looked up.
template <class M, class E>
__device__ __forceinline__ void rk_end_stage(const DPot& P, double dir,
                                             double te, const double (&q1)[3],
                                             double (&g)[M::SF][3]) {
  if constexpr (M::SF > M::S) {
    constexpr int l = M::SF - 1;
    if constexpr (track_prefetch<E>::value)
      E::acc_at(P, q1[0], q1[1], q1[2], dir * te, g[l][0], g[l][1], g[l][2],
                g[l][0], g[l][1], g[l][2]);
    else
      E::acc(P, q1[0], q1[1], q1[2], dir * te, g[l][0], g[l][1], g[l][2]);
  }
}


template <int D>
__device__ __forceinline__ double horner(const double (&c)[D], double x) {
  double acc = c[0];
#pragma unroll
  for (int d = 1; d < D; ++d) acc = fma(acc, x, c[d]);
  return acc;
}

// The dense output of a step (q, p at its start, step h, stage
// accelerations g[0..SF-1] including the end stage) at th in [0, 1]:
// diffrax's interpolant for the method, in Runge-Kutta-Nystrom form.
// (The coefficient rows go to horner() by reference: the compiler then
// loads them from the constant table into scalar registers once, which
// measured lighter on the register file than folding them into the code.)
template <class M>
__device__ __forceinline__ void dense_eval(double th, double h,
                                           const double (&q)[3],
                                           const double (&p)[3],
                                           const double (&g)[M::SF][3],
                                           double (&qo)[3], double (&po)[3]) {
  constexpr int SF = M::SF;
  double sq0 = 0.0, sq1 = 0.0, sq2 = 0.0, sp0 = 0.0, sp1 = 0.0, sp2 = 0.0;
#pragma unroll
  for (int l = 0; l < SF; ++l) {
    if (M::dn.nq[l]) {
      const double w = horner(M::dn.wq[l], th);
      sq0 = fma(w, g[l][0], sq0);
      sq1 = fma(w, g[l][1], sq1);
      sq2 = fma(w, g[l][2], sq2);
    }
    if (M::dn.np[l]) {
      const double w = horner(M::dn.wp[l], th);
      sp0 = fma(w, g[l][0], sp0);
      sp1 = fma(w, g[l][1], sp1);
      sp2 = fma(w, g[l][2], sp2);
    }
  }
  const double ht = h * th, hht = h * ht;
  po[0] = fma(ht, sp0, p[0]);
  po[1] = fma(ht, sp1, p[1]);
  po[2] = fma(ht, sp2, p[2]);
  qo[0] = fma(hht, sq0, fma(ht, p[0], q[0]));
  qo[1] = fma(hht, sq1, fma(ht, p[1], q[1]));
  qo[2] = fma(hht, sq2, fma(ht, p[2], q[2]));
}

// One accepted step's dense-output record, in physical variables: start
// time, signed step, q, v, and the SF stage accelerations (the dense output
// over the step is q(th) = q + th h v + h^2 th sum Wq g, v(th) = v + h th
// sum Wp g with h the signed step; see gx_dense_eval).
template <class M>
__device__ __forceinline__ void write_record(double* r, double dir, double tau,
                                             double h, const double (&q)[3],
                                             const double (&p)[3],
                                             const double (&g)[M::SF][3]) {
  r[0] = dir * tau;
  r[1] = dir * h;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    r[2 + m] = q[m];
    r[5 + m] = dir * p[m];
  }
#pragma unroll
  for (int l = 0; l < M::SF; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m) r[8 + 3 * l + m] = g[l][m];
}

// The dense output's weights at th, for callers that combine stages held in
// memory: q(th) = q + th h p + h^2 th sum_l wq_l g_l and
// p(th) = p + h th sum_l wp_l g_l.
template <class M>
__device__ __forceinline__ void dense_weights(double th, double (&wq)[M::SF],
                                              double (&wp)[M::SF]) {
  constexpr int D = sizeof(M::dn.wq[0]) / sizeof(double);
#pragma unroll
  for (int l = 0; l < M::SF; ++l) {
    double a = M::dn.wq[l][0], b = M::dn.wp[l][0];
#pragma unroll
    for (int d = 1; d < D; ++d) {
      a = fma(a, th, M::dn.wq[l][d]);
      b = fma(b, th, M::dn.wp[l][d]);
    }
    wq[l] = a;
    wp[l] = b;
  }
}

}  // namespace gx
