// gx_rk.hpp -- one evaluation of an explicit Runge-Kutta-Nystrom method
// over one orbit, shared by the independent-orbit and the shared-step
// integrators.
#pragma once

#include "gx_potential.hpp"
#include "gx_tableau.hpp"

namespace gx {

// One evaluation of method M from (tb, q, p) with step h (solver time tau;
// physical time dir * tau, velocities pre-multiplied by dir).
//
// The stage loop is kept rolled so that the (large) potential evaluation is
// emitted once; each stage's coefficients and its slot in g[][] are selected
// by a chain of compile-time-indexed branches on the (wave-uniform) stage
// number, which keeps g[][] in registers.
template <class M, class E, bool ERR>
__device__ __forceinline__ void rk_eval(const DPot& P, double dir, double tb,
                                        double h, const double (&q)[3],
                                        const double (&p)[3], double (&q1)[3],
                                        double (&p1)[3], double (&eq)[3],
                                        double (&ep)[3]) {
  constexpr int S = M::S;
  double g[S][3];
  // Stage positions as q + rs (h p) + (h h) s: two FMAs per component
  const double hh = h * h;
  const double hp[3] = {h * p[0], h * p[1], h * p[2]};
  if constexpr (!E::kLarge) {
    // Small evaluator: unroll the stages; every index below is static.
    E::acc(P, q[0], q[1], q[2], dir * tb, g[0][0], g[0][1], g[0][2]);
#pragma unroll
    for (int i = 1; i < S; ++i) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      bool any = false;
#pragma unroll
      for (int l = 0; l < i; ++l) {
        if (M::k.aa[i][l] != 0.0) {
          const double a = M::k.aa[i][l];
          s0 = fma(a, g[l][0], s0);
          s1 = fma(a, g[l][1], s1);
          s2 = fma(a, g[l][2], s2);
          any = true;
        }
      }
      const double rs = M::k.rs[i];
      double Q0 = fma(rs, hp[0], q[0]), Q1 = fma(rs, hp[1], q[1]),
             Q2 = fma(rs, hp[2], q[2]);
      if (any) {
        Q0 = fma(hh, s0, Q0);
        Q1 = fma(hh, s1, Q1);
        Q2 = fma(hh, s2, Q2);
      }
      E::acc(P, Q0, Q1, Q2, dir * fma(M::k.c[i], h, tb), g[i][0], g[i][1],
             g[i][2]);
    }
  } else {
    // Large evaluator: keep the stage loop rolled so that it is emitted
    // once; the stage's coefficients and its slot in g[][] are selected by
    // real scalar branches on the wave-uniform stage number.
#pragma unroll
    for (int k = 0; k < S; ++k) g[k][0] = g[k][1] = g[k][2] = 0.0;
    double Qx = q[0], Qy = q[1], Qz = q[2], ts = tb;
#pragma nounroll
    for (int i = 0; i < S; ++i) {
      if (i > 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, rs = 0.0, c = 0.0;
#pragma unroll
        for (int k = 1; k < S; ++k) {
          if (k == i) {
            asm volatile("");  // no if-conversion into selects
#pragma unroll
            for (int l = 0; l < k; ++l) {
              const double a = M::k.aa[k][l];
              if (a != 0.0) {
                s0 = fma(a, g[l][0], s0);
                s1 = fma(a, g[l][1], s1);
                s2 = fma(a, g[l][2], s2);
              }
            }
            rs = M::k.rs[k];
            c = M::k.c[k];
          }
        }
        Qx = fma(hh, s0, fma(rs, hp[0], q[0]));
        Qy = fma(hh, s1, fma(rs, hp[1], q[1]));
        Qz = fma(hh, s2, fma(rs, hp[2], q[2]));
        ts = fma(c, h, tb);
      }
      double ax, ay, az;
      E::acc(P, Qx, Qy, Qz, dir * ts, ax, ay, az);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if (k == i) {
          asm volatile("");
          g[k][0] = ax;
          g[k][1] = ay;
          g[k][2] = az;
        }
      }
    }
  }
  double b0 = 0.0, b1 = 0.0, b2 = 0.0, ba0 = 0.0, ba1 = 0.0, ba2 = 0.0;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0, ea0 = 0.0, ea1 = 0.0, ea2 = 0.0;
#pragma unroll
  for (int l = 0; l < S; ++l) {
    if (M::k.b[l] != 0.0) {
      const double b = M::k.b[l];
      b0 = fma(b, g[l][0], b0);
      b1 = fma(b, g[l][1], b1);
      b2 = fma(b, g[l][2], b2);
    }
    if (M::k.ba[l] != 0.0) {
      const double ba = M::k.ba[l];
      ba0 = fma(ba, g[l][0], ba0);
      ba1 = fma(ba, g[l][1], ba1);
      ba2 = fma(ba, g[l][2], ba2);
    }
    if (ERR) {
      if (M::k.e[l] != 0.0) {
        const double e = M::k.e[l];
        e0 = fma(e, g[l][0], e0);
        e1 = fma(e, g[l][1], e1);
        e2 = fma(e, g[l][2], e2);
      }
      if (M::k.ea[l] != 0.0) {
        const double ea = M::k.ea[l];
        ea0 = fma(ea, g[l][0], ea0);
        ea1 = fma(ea, g[l][1], ea1);
        ea2 = fma(ea, g[l][2], ea2);
      }
    }
  }
  constexpr double sb = M::k.sb, se = M::k.se;
  p1[0] = fma(h, b0, p[0]);
  p1[1] = fma(h, b1, p[1]);
  p1[2] = fma(h, b2, p[2]);
  q1[0] = fma(hh, ba0, fma(sb, hp[0], q[0]));
  q1[1] = fma(hh, ba1, fma(sb, hp[1], q[1]));
  q1[2] = fma(hh, ba2, fma(sb, hp[2], q[2]));
  if (ERR) {
    ep[0] = h * e0;
    ep[1] = h * e1;
    ep[2] = h * e2;
    eq[0] = fma(hh, ea0, se * hp[0]);
    eq[1] = fma(hh, ea1, se * hp[1]);
    eq[2] = fma(hh, ea2, se * hp[2]);
  }
}

}  // namespace gx
