// gx_tableau.hpp -- Butcher tableaux of the two explicit Runge-Kutta pairs
// galax integrates with (diffrax Dopri8 / Dopri5, called from
// dynamics/_src/orbit/solver.py:121-141 and
// dynamics/_src/experimental/stream.py:32-39), and their Runge-Kutta-Nystrom
// reduction for Hamilton's equations.
//
// For y = (q, p) with dq/dt = p and dp/dt = a(q, t), a stage of the standard
// scheme is Y_i = y + h sum_j a_ij f_j.  Because f_j = (P_j, a(Q_j)) and
// P_j = p + h sum_l a_jl a(Q_l), the position stages and the step outputs can
// be written with the accelerations alone:
//   Q_i  = q + h (rs_i p + h sum_l (A A)_il g_l),   rs_i = sum_j a_ij
//   q1   = q + h (sum_i b_i  p + h sum_l (b A)_l  g_l)
//   p1   = p + h sum_i b_i g_i
//   err_q = h (sum_i e_i p + h sum_l (e A)_l g_l),   err_p = h sum_i e_i g_i
// with g_l = a(Q_l, t + c_l h).  This is the same arithmetic as the standard
// scheme up to rounding, needs no velocity stages (a does not depend on p),
// and keeps 3 numbers per stage instead of 6 in registers.  All derived
// coefficients are computed here at compile time in IEEE double.
#pragma once

namespace gx {

template <int S>
struct Tab {
  double c[S];
  double a[S][S];
  double b[S];
  double e[S];  // b_sol - b_embedded
};

template <int S>
struct Rkn {
  double c[S];
  double a[S][S];   // for p-derivative uses (adjoint)
  double rs[S];     // row sums of a
  double aa[S][S];  // A*A
  double b[S];
  double ba[S];     // b*A
  double e[S];
  double ea[S];     // e*A
  double sb, se;    // sum b, sum e
};

template <int S>
constexpr Rkn<S> make_rkn(const Tab<S>& t) {
  Rkn<S> r{};
  for (int i = 0; i < S; ++i) {
    r.c[i] = t.c[i];
    r.b[i] = t.b[i];
    r.e[i] = t.e[i];
    double s = 0;
    for (int j = 0; j < S; ++j) {
      r.a[i][j] = t.a[i][j];
      s += t.a[i][j];
    }
    r.rs[i] = s;
  }
  for (int i = 0; i < S; ++i)
    for (int l = 0; l < S; ++l) {
      double s = 0;
      for (int j = 0; j < S; ++j) s += t.a[i][j] * t.a[j][l];
      r.aa[i][l] = s;
    }
  double sb = 0, se = 0;
  for (int l = 0; l < S; ++l) {
    double x = 0, y = 0;
    for (int i = 0; i < S; ++i) {
      x += t.b[i] * t.a[i][l];
      y += t.e[i] * t.a[i][l];
    }
    r.ba[l] = x;
    r.ea[l] = y;
    sb += t.b[l];
    se += t.e[l];
  }
  r.sb = sb;
  r.se = se;
  return r;
}

// Prince & Dormand (1981) RK8(7)13M.
constexpr Tab<13> dopri8_tab() {
  Tab<13> t{};
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
  for (int i = 0; i < 13; ++i) t.c[i] = c[i];
  t.a[1][0] = 1.0 / 18;
  t.a[2][0] = 1.0 / 48;
  t.a[2][1] = 1.0 / 16;
  t.a[3][0] = 1.0 / 32;
  t.a[3][2] = 3.0 / 32;
  t.a[4][0] = 5.0 / 16;
  t.a[4][2] = -75.0 / 64;
  t.a[4][3] = 75.0 / 64;
  t.a[5][0] = 3.0 / 80;
  t.a[5][3] = 3.0 / 16;
  t.a[5][4] = 3.0 / 20;
  t.a[6][0] = 29443841.0 / 614563906.0;
  t.a[6][3] = 77736538.0 / 692538347.0;
  t.a[6][4] = -28693883.0 / 1125000000.0;
  t.a[6][5] = 23124283.0 / 1800000000.0;
  t.a[7][0] = 16016141.0 / 946692911.0;
  t.a[7][3] = 61564180.0 / 158732637.0;
  t.a[7][4] = 22789713.0 / 633445777.0;
  t.a[7][5] = 545815736.0 / 2771057229.0;
  t.a[7][6] = -180193667.0 / 1043307555.0;
  t.a[8][0] = 39632708.0 / 573591083.0;
  t.a[8][3] = -433636366.0 / 683701615.0;
  t.a[8][4] = -421739975.0 / 2616292301.0;
  t.a[8][5] = 100302831.0 / 723423059.0;
  t.a[8][6] = 790204164.0 / 839813087.0;
  t.a[8][7] = 800635310.0 / 3783071287.0;
  t.a[9][0] = 246121993.0 / 1340847787.0;
  t.a[9][3] = -37695042795.0 / 15268766246.0;
  t.a[9][4] = -309121744.0 / 1061227803.0;
  t.a[9][5] = -12992083.0 / 490766935.0;
  t.a[9][6] = 6005943493.0 / 2108947869.0;
  t.a[9][7] = 393006217.0 / 1396673457.0;
  t.a[9][8] = 123872331.0 / 1001029789.0;
  t.a[10][0] = -1028468189.0 / 846180014.0;
  t.a[10][3] = 8478235783.0 / 508512852.0;
  t.a[10][4] = 1311729495.0 / 1432422823.0;
  t.a[10][5] = -10304129995.0 / 1701304382.0;
  t.a[10][6] = -48777925059.0 / 3047939560.0;
  t.a[10][7] = 15336726248.0 / 1032824649.0;
  t.a[10][8] = -45442868181.0 / 3398467696.0;
  t.a[10][9] = 3065993473.0 / 597172653.0;
  t.a[11][0] = 185892177.0 / 718116043.0;
  t.a[11][3] = -3185094517.0 / 667107341.0;
  t.a[11][4] = -477755414.0 / 1098053517.0;
  t.a[11][5] = -703635378.0 / 230739211.0;
  t.a[11][6] = 5731566787.0 / 1027545527.0;
  t.a[11][7] = 5232866602.0 / 850066563.0;
  t.a[11][8] = -4093664535.0 / 808688257.0;
  t.a[11][9] = 3962137247.0 / 1805957418.0;
  t.a[11][10] = 65686358.0 / 487910083.0;
  t.a[12][0] = 403863854.0 / 491063109.0;
  t.a[12][3] = -5068492393.0 / 434740067.0;
  t.a[12][4] = -411421997.0 / 543043805.0;
  t.a[12][5] = 652783627.0 / 914296604.0;
  t.a[12][6] = 11173962825.0 / 925320556.0;
  t.a[12][7] = -13158990841.0 / 6184727034.0;
  t.a[12][8] = 3936647629.0 / 1978049680.0;
  t.a[12][9] = -160528059.0 / 685178525.0;
  t.a[12][10] = 248638103.0 / 1413531060.0;
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
    t.b[i] = b[i];
    t.e[i] = b[i] - bh[i];
  }
  return t;
}

// Dormand & Prince (1980) RK5(4)7M with diffrax's embedded weights.
constexpr Tab<7> dopri5_tab() {
  Tab<7> t{};
  const double c[7] = {0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
  for (int i = 0; i < 7; ++i) t.c[i] = c[i];
  t.a[1][0] = 1.0 / 5;
  t.a[2][0] = 3.0 / 40;
  t.a[2][1] = 9.0 / 40;
  t.a[3][0] = 44.0 / 45;
  t.a[3][1] = -56.0 / 15;
  t.a[3][2] = 32.0 / 9;
  t.a[4][0] = 19372.0 / 6561;
  t.a[4][1] = -25360.0 / 2187;
  t.a[4][2] = 64448.0 / 6561;
  t.a[4][3] = -212.0 / 729;
  t.a[5][0] = 9017.0 / 3168;
  t.a[5][1] = -355.0 / 33;
  t.a[5][2] = 46732.0 / 5247;
  t.a[5][3] = 49.0 / 176;
  t.a[5][4] = -5103.0 / 18656;
  t.a[6][0] = 35.0 / 384;
  t.a[6][2] = 500.0 / 1113;
  t.a[6][3] = 125.0 / 192;
  t.a[6][4] = -2187.0 / 6784;
  t.a[6][5] = 11.0 / 84;
  const double b[7] = {35.0 / 384,     0,          500.0 / 1113, 125.0 / 192,
                       -2187.0 / 6784, 11.0 / 84,  0};
  const double e[7] = {35.0 / 384 - 1951.0 / 21600,
                       0,
                       500.0 / 1113 - 22642.0 / 50085,
                       125.0 / 192 - 451.0 / 720,
                       -2187.0 / 6784 + 12231.0 / 42400,
                       11.0 / 84 - 649.0 / 6300,
                       -1.0 / 60.0};
  for (int i = 0; i < 7; ++i) {
    t.b[i] = b[i];
    t.e[i] = e[i];
  }
  return t;
}

// The tableau with one more stage at (t0 + h, y1): row b, c = 1, no weight
// in the solution or the error.  Its acceleration a(q1, t1) is the next
// step's first stage (first same as last) and the 14th increment that
// Dopri8's dense output reads.
template <int S>
constexpr Tab<S + 1> with_end_stage(const Tab<S>& t) {
  Tab<S + 1> r{};
  for (int i = 0; i < S; ++i) {
    r.c[i] = t.c[i];
    r.b[i] = t.b[i];
    r.e[i] = t.e[i];
    for (int j = 0; j < S; ++j) r.a[i][j] = t.a[i][j];
  }
  r.c[S] = 1.0;
  for (int j = 0; j < S; ++j) r.a[S][j] = t.b[j];
  return r;
}

// Dense output over the SF stage slots of a step whose last slot is the
// end stage, in Runge-Kutta-Nystrom form.  With first-order weights w_l(th)
// (y(t0 + th h) = y0 + sum_l w_l(th) h f_l, sum_l w_l = th):
//   p(th) = p + h th sum_l Wp_l(th) g_l,          Wp_l = w_l / th
//   q(th) = q + th (h p) + h^2 th sum_l Wq_l(th) g_l,
//                                     Wq_l = sum_i Wp_i a_il
// because the velocity stage i is p + h sum_l a_il g_l.  Polynomials are
// stored in Horner order (th^(D-1) first).
template <int SF, int D>
struct Dense {
  double wp[SF][D];
  double wq[SF][D];
  bool np[SF], nq[SF];  // rows that are not identically zero
};

template <int SF, int D>
constexpr Dense<SF, D> make_dense(const double (&wp)[SF][D],
                                  const Tab<SF>& t) {
  Dense<SF, D> r{};
  for (int l = 0; l < SF; ++l)
    for (int d = 0; d < D; ++d) {
      r.wp[l][d] = wp[l][d];
      double s = 0.0;
      for (int i = 0; i < SF; ++i) s += wp[i][d] * t.a[i][l];
      r.wq[l][d] = s;
    }
  for (int l = 0; l < SF; ++l) {
    bool a = false, b = false;
    for (int d = 0; d < D; ++d) {
      a = a || r.wp[l][d] != 0.0;
      b = b || r.wq[l][d] != 0.0;
    }
    r.np[l] = a;
    r.nq[l] = b;
  }
  return r;
}

// Dopri8's interpolant (diffrax 0.7.0 Dopri8 dense output, a continuous
// extension of RK8(7)13M over its 13 stages and the end stage): w_l(th) =
// th * (quintic in th).  Stages 2-5 carry no weight; stage 1's row follows
// from sum_l w_l = th.
constexpr Dense<14, 6> dopri8_dense() {
  double w[14][6] = {};
  const double rows[9][6] = {
      {-39.6107919852202505218, 116.4422149550342161651,
       -121.4999627731334642623, 52.2273532792945524050,
       -7.6142658045872677172, 0},
      {20.3761213808791436958, -67.1451318825957197185,
       83.1721004639847717481, -46.8919164181093621583,
       10.7281392630428866124, 0},
      {7.3347098826795362023, -16.5672243527496524646, 9.5724507555993664382,
       -0.1890893225010595467, 0.5526637063753648783, 0},
      {32.8801774352459155182, -89.9916014847245016028,
       87.8406057677205645007, -35.7075975946222072821,
       4.2186562625665153803, 0},
      {-10.1588990526426760954, 22.6237489648532849093,
       -17.4152107770762969005, 6.2736448083240352160,
       -0.6627209125361597559, 0},
      {-12.5401268098782561200, 32.2362340167355370113,
       -28.5903289514790976966, 10.3160881272450748458,
       -1.2636789001135462218, 0},
      {29.5553001484516038033, -82.1020315488359848644,
       81.6630950584341412934, -34.7650769866611817349,
       5.4106037898590422230, 0},
      {-41.7923486424390588923, 116.2662185791119533462,
       -114.9375291377009418170, 47.7457971078225540396,
       -7.0321379067945741781, 0},
      {20.3006925822100825485, -53.9020777466385396792,
       50.2558364226176017553, -19.0082099341608028453,
       2.3537586759714983486, 0}};
  for (int r = 0; r < 9; ++r)
    for (int d = 0; d < 6; ++d) w[5 + r][d] = rows[r][d];
  for (int d = 0; d < 6; ++d) {
    double s = (d == 5) ? 1.0 : 0.0;
    for (int i = 5; i < 14; ++i) s -= w[i][d];
    w[0][d] = s;
  }
  return make_dense<14, 6>(w, with_end_stage(dopri8_tab()));
}

// Dopri5's interpolant (diffrax's, and the one scipy's RK45 uses): the
// quartic through y0, y1, the end slopes f0 = k_1, f1 = k_7 and Shampine's
// (1986) midpoint value y0 + sum_i m_i k_i.  Written over the increments,
// w_i(th) / th = a_i th^3 + b_i th^2 + c_i th + [i = 1] with
//   a_i = 2([i=7] - [i=1]) - 8 b_i + 16 m_i
//   b_i = 5 [i=1] - 3 [i=7] + 14 b_i - 32 m_i
//   c_i = [i=7] - 4 [i=1] - 5 b_i + 16 m_i
// (the y0 terms of the quartic cancel).  Its 7th stage is already at y1.
constexpr Dense<7, 4> dopri5_dense() {
  const Tab<7> t = dopri5_tab();
  const double m[7] = {6025192743.0 / 30085553152.0 / 2,
                       0,
                       51252292925.0 / 65400821598.0 / 2,
                       -2691868925.0 / 45128329728.0 / 2,
                       187940372067.0 / 1594534317056.0 / 2,
                       -1776094331.0 / 19743644256.0 / 2,
                       11237099.0 / 235043384.0 / 2};
  double w[7][4] = {};
  for (int i = 0; i < 7; ++i) {
    const double first = (i == 0) ? 1.0 : 0.0, last = (i == 6) ? 1.0 : 0.0;
    w[i][0] = 2.0 * (last - first) - 8.0 * t.b[i] + 16.0 * m[i];
    w[i][1] = 5.0 * first - 3.0 * last + 14.0 * t.b[i] - 32.0 * m[i];
    w[i][2] = last - 4.0 * first - 5.0 * t.b[i] + 16.0 * m[i];
    w[i][3] = first;
  }
  return make_dense<7, 4>(w, t);
}

// S: stages of the method; SF: stage slots of a step including the end
// stage (Dopri5's last stage is its end stage); kf: the Runge-Kutta-Nystrom
// form over the SF slots; dn: the dense output over them.
struct Dopri8 {
  static constexpr int S = 13;
  static constexpr int SF = 14;
  static constexpr int order = 8;
  static constexpr Rkn<13> k = make_rkn(dopri8_tab());
  static constexpr Rkn<14> kf = make_rkn(with_end_stage(dopri8_tab()));
  static constexpr Dense<14, 6> dn = dopri8_dense();
};
struct Dopri5 {
  static constexpr int S = 7;
  static constexpr int SF = 7;
  static constexpr int order = 5;
  static constexpr Rkn<7> k = make_rkn(dopri5_tab());
  static constexpr Rkn<7> kf = make_rkn(dopri5_tab());
  static constexpr Dense<7, 4> dn = dopri5_dense();
};

}  // namespace gx
