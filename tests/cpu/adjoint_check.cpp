// CPU check of the reverse-mode primitives of the tracking kernels
// (accel-plus_amd/csrc/accp_adjoint.h, compiled here for the host): for each
// element map, (1) the inverse recovers the input state, (2) the
// vector-Jacobian product equals central finite differences of the forward
// map, including the knob derivatives; (3) the forms that carry several
// adjoint vectors give, vector by vector, what the one-vector form gives.
// Prints one line per check and exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

#include "accp_adjoint.h"

using namespace accp;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static double inv_fact(int j) {
  long double f = 1.0L;
  for (int i = 2; i <= j; ++i) f *= i;
  return (double)(1.0L / f);
}
static std::vector<double> quad_coefs(int nt) {
  std::vector<double> P;
  const int M = nt / 2, sh[4] = {0, 2, 1, 3};
  for (int b = 0; b < 4; ++b)
    for (int m = 0; m < M; ++m) P.push_back(inv_fact(4 * m + sh[b]));
  for (int m = 0; m < M; ++m) P.push_back((2.0 * m + 1) * inv_fact(4 * m + 3));
  for (int m = 0; m < M; ++m) P.push_back((2.0 * m + 2) * inv_fact(4 * m + 5));
  return P;
}

static Part mk(const double v[6]) {
  Part p{v[0], v[1], v[2], v[3], v[4], 0, 0};
  set_delta(p, v[5]);
  return p;
}
static void unpack(const Part &p, double v[6]) {
  v[0] = p.x; v[1] = p.px; v[2] = p.y; v[3] = p.py; v[4] = p.z; v[5] = p.d;
}

static int failures = 0;

// fwd(state, knob) -> state; rev(state_out, adj, grad) -> state_in, adj_in
// knob_rel: relative finite-difference step on the knob (maps linear in
// their knob take a large one, which is exact for them); coord_rel: the same
// for the coordinates (strongly focusing maps amplify round-off and take a
// larger one)
static void check(const char *name, std::function<void(Part &, double)> fwd,
                  std::function<void(Part &, Adj &, double *)> rev, double knob, bool has_knob,
                  const double in[6], double knob_rel = 1e-7, double tol = 2e-7, double coord_rel = 1e-6) {
  // forward and inverse round trip
  Part p = mk(in);
  fwd(p, knob);
  double out[6];
  unpack(p, out);
  const double lw[6] = {U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
  Adj l{lw[0], lw[1], lw[2], lw[3], lw[4], lw[5]};
  double grad[4] = {0, 0, 0, 0};
  rev(p, l, grad);
  double back[6];
  unpack(p, back);
  double rt = 0;
  // delta comes back through 1 + delta in the maps that move it: its floor
  // is 1e-3 (an ulp of 1 is then 2e-13 of it)
  for (int i = 0; i < 6; ++i)
    rt = std::fmax(rt, std::fabs(back[i] - in[i]) / (std::fabs(in[i]) + (i == 5 ? 1e-3 : 1e-6)));
  // VJP against central differences
  const double lin[6] = {l.x, l.px, l.y, l.py, l.z, l.d};
  double worst = 0;
  for (int i = 0; i <= 6; ++i) {
    if (i == 6 && !has_knob) break;
    const double hstep = i < 6 ? coord_rel * (std::fabs(in[i]) + 1e-4) : knob_rel * (std::fabs(knob) + 1e-3);
    double vp[6], vm[6];
    for (int k = 0; k < 6; ++k) vp[k] = vm[k] = in[k];
    double kp = knob, km = knob;
    if (i < 6) { vp[i] += hstep; vm[i] -= hstep; } else { kp += hstep; km -= hstep; }
    Part a = mk(vp), b = mk(vm);
    fwd(a, kp);
    fwd(b, km);
    double oa[6], ob[6];
    unpack(a, oa);
    unpack(b, ob);
    double fd = 0;
    for (int k = 0; k < 6; ++k) fd += lw[k] * (oa[k] - ob[k]) / (2 * hstep);
    const double ad = i < 6 ? lin[i] : grad[0];
    const double err = std::fabs(fd - ad) / (std::fabs(fd) + std::fabs(ad) + 1e-9);
    worst = std::fmax(worst, err);
  }
  const bool ok = rt < 1e-11 && worst < tol;
  std::printf("%-28s roundtrip %.2e  vjp-vs-fd %.2e  %s\n", name, rt, worst, ok ? "ok" : "FAIL");
  if (!ok) ++failures;
}

// The forms that carry several adjoint vectors (*_n<NV>) against the
// one-vector form run once per vector from the same state: the same
// inverse state and, vector by vector, the same adjoint and knob
// derivative, bit for bit (the host build contracts nothing).
struct GradRows {
  double (*g)[4];
  void operator()(int v, int slot, double x) const { g[v][slot] += x; }
};
static void check_vec(const char *name, std::function<void(Part &, Adj (&)[3], GradRows)> rev3,
                      std::function<void(Part &, Adj &, double *)> rev1, const double in[6]) {
  Adj l3[3];
  double w[3][6];
  for (int v = 0; v < 3; ++v) {
    for (int k = 0; k < 6; ++k) w[v][k] = U(-1, 1);
    l3[v] = Adj{w[v][0], w[v][1], w[v][2], w[v][3], w[v][4], w[v][5]};
  }
  double g3[3][4] = {};
  Part p3 = mk(in);
  rev3(p3, l3, GradRows{g3});
  bool same = true;
  for (int v = 0; v < 3; ++v) {
    Part p1 = mk(in);
    Adj l1{w[v][0], w[v][1], w[v][2], w[v][3], w[v][4], w[v][5]};
    double g1[4] = {};
    rev1(p1, l1, g1);
    const double a[13] = {p1.x, p1.px, p1.y, p1.py, p1.z, p1.d, l1.x, l1.px, l1.y, l1.py, l1.z, l1.d, g1[1]};
    const double b[13] = {p3.x, p3.px, p3.y, p3.py, p3.z, p3.d, l3[v].x, l3[v].px, l3[v].y, l3[v].py,
                          l3[v].z, l3[v].d, g3[v][1]};
    for (int k = 0; k < 13; ++k) same = same && (a[k] == b[k] || (a[k] != a[k] && b[k] != b[k]));
  }
  std::printf("%-28s three vectors = one vector three times: %s\n", name, same ? "ok" : "FAIL");
  if (!same) ++failures;
}

static void vector_forms(const double s0[6]) {
  check_vec("drift x3", [](Part &p, Adj (&l)[3], GradRows) { drift_nl_rev_n<3>(p, l, 0.7); },
            [](Part &p, Adj &l, double *) { drift_nl_rev(p, l, 0.7); }, s0);
  check_vec("quad K=0 x3", [](Part &p, Adj (&l)[3], GradRows) { drift_kzero_rev_n<3>(p, l, 0.3); },
            [](Part &p, Adj &l, double *) { drift_kzero_rev(p, l, 0.3); }, s0);
  check_vec("edge x3", [](Part &p, Adj (&l)[3], GradRows) { edge_rev_n<3>(p, l, 0.013); },
            [](Part &p, Adj &l, double *) { edge_rev(p, l, 0.013); }, s0);
  const double h = 0.157 / 1.5, L = 0.15 * DRIFT1, hs = h * L, hq = 1 / 4. * hs * L;
  check_vec("hxdrift x3",
            [=](Part &p, Adj (&l)[3], GradRows) { hxdrift_c_rev_n<3>(p, l, hx_consts(p.ipd, L, hs, hq), h, hs * 0.5); },
            [=](Part &p, Adj &l, double *) { hxdrift_c_rev(p, l, hx_consts(p.ipd, L, hs, hq), h, hs * 0.5); }, s0);
  const double kl = 0.15 * KICK1, lh = kl * h;
  const double b0[1] = {2e-4}, b2[3] = {1e-4, -0.3, 4.0};
  check_vec("bend kick MO=0 x3",
            [=](Part &p, Adj (&l)[3], GradRows) { bend_kick_rev_n<3, 0, true>(p, l, b0, kl, h, lh * p.d, lh); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<0, true>(p, l, b0, kl, h, lh * p.d, lh); }, s0);
  check_vec("bend kick MO=2 x3",
            [=](Part &p, Adj (&l)[3], GradRows) {
              bend_kick_rev_n<3, 2, false>(p, l, b2, kl, h, lh * p.d, lh, -0.3, 4.0);
            },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<2, false>(p, l, b2, kl, h, lh * p.d, lh, -0.3, 4.0); },
            s0);
  double b9[10], a9[10];
  for (int i = 0; i < 10; ++i) {
    b9[i] = U(-1, 1) * std::pow(10.0, i);
    a9[i] = U(-1, 1) * std::pow(10.0, i);
  }
  for (int kind = 2; kind <= 3; ++kind) {
    const Knob kn{1, kind, 4};
    check_vec(kind == 2 ? "multipole kick, PolyB knob x3" : "multipole kick, PolyA knob x3",
              [=](Part &p, Adj (&l)[3], GradRows g) { mpole_kick_rev_n<3, 9>(p, l, b9, a9, 0.02, kn, g); },
              [=](Part &p, Adj &l, double *g) { mpole_kick_rev<9>(p, l, b9, a9, 0.02, kn, g); }, s0);
    check_vec(kind == 2 ? "empty multipole, PolyB knob x3" : "empty multipole, PolyA knob x3",
              [=](Part &p, Adj (&l)[3], GradRows g) { mpole_knob_only_n<3>(p, l, 0.02, kn, g); },
              [=](Part &p, Adj &l, double *g) { mpole_knob_only(p, l, 0.02, kn, g); }, s0);
  }
  static const std::vector<double> C8 = quad_coefs(8);
  for (double K : {1.2, -1.1}) {
    check_vec(K > 0 ? "series quad K>0 x3" : "series quad K<0 x3",
              [=](Part &p, Adj (&l)[3], GradRows g) {
                quad_series_rev_n<3, 8>(p, l, C8.data(), 0.3, 0.09, K > 0, std::fabs(K), K > 0 ? 1.0 : -1.0,
                                        Knob{1, 1, 0}, g);
              },
              [=](Part &p, Adj &l, double *g) {
                quad_series_rev<8>(p, l, C8.data(), 0.3, 0.09, K > 0, std::fabs(K), K > 0 ? 1.0 : -1.0,
                                   Knob{1, 1, 0}, g);
              },
              s0);
  }
  const double wc = TWOPI * 562e6 / C0;
  check_vec("cavity x3", [=](Part &p, Adj (&l)[3], GradRows) { cavity_rev_n<3>(p, l, 0.4, 5e-4, wc, 0.1, true); },
            [=](Part &p, Adj &l, double *) { cavity_rev(p, l, 0.4, 5e-4, wc, 0.1, true); }, s0);
  check_vec("corrector x3", [](Part &p, Adj (&l)[3], GradRows) { corrector_rev_n<3>(p, l, 0.2, 1e-4, -2e-4); },
            [](Part &p, Adj &l, double *) { corrector_rev(p, l, 0.2, 1e-4, -2e-4); }, s0);
}

int main() {
  for (int trial = 0; trial < 3; ++trial) {
    const double d = trial == 0 ? -0.03 : trial == 1 ? 0.0 : 0.027;
    const double s0[6] = {U(-3e-3, 3e-3), U(-3e-4, 3e-4), U(-2e-3, 2e-3), U(-2e-4, 2e-4), U(-1e-3, 1e-3), d};

    check("drift", [](Part &p, double) { drift_nl(p, 0.7 * p.ipd); },
          [](Part &p, Adj &l, double *) { drift_nl_rev(p, l, 0.7); }, 0, false, s0);

    check("quad K=0 (drift quirk)", [](Part &p, double) { drift_kzero(p, 0.3); },
          [](Part &p, Adj &l, double *) { drift_kzero_rev(p, l, 0.3); }, 0, false, s0);

    {  // a coupled 6x6 matrix between half drifts (near the identity, det != 1)
      static double M[36];
      for (int i = 0; i < 36; ++i) M[i] = (i % 7 == 0 ? 1.0 : 0.0) + 0.05 * std::sin(1.3 * i + 0.4);
      check("matrix", [](Part &p, double) { matrix_map(p, 0.4, M); },
            [](Part &p, Adj &l, double *) { matrix_rev(p, l, 0.4, M); }, 0, false, s0);
    }

    check("edge", [](Part &p, double) { edge(p, 0.013); }, [](Part &p, Adj &l, double *) { edge_rev(p, l, 0.013); },
          0, false, s0);

    {  // pseudo-drift of a bend
      const double h = 0.157 / 1.5, L = 0.15 * DRIFT1, hs = h * L, hq = 1 / 4. * hs * L;
      check("hxdrift", [=](Part &p, double) { hxdrift_c(p, hx_consts(p.ipd, L, hs, hq), h, hs * 0.5); },
            [=](Part &p, Adj &l, double *) { hxdrift_c_rev(p, l, hx_consts(p.ipd, L, hs, hq), h, hs * 0.5); }, 0,
            false, s0);
    }
    {  // bend kicks, dipole and with normal multipoles
      const double h = 0.157 / 1.5, L = 0.15 * KICK1, lh = L * h;
      const double b0[1] = {0.0};
      check("bend kick MO=0", [=](Part &p, double) { bend_kick_c<0, true>(p, b0, L, h, lh * p.d, lh, 0, 0); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<0, true>(p, l, b0, L, h, lh * p.d, lh); }, 0, false,
            s0);
      const double b3[4] = {1e-3, 0.0, 5.0, 20.0};
      check("bend kick MO=3", [=](Part &p, double) { bend_kick_c<3, true>(p, b3, L, h, lh * p.d, lh, 0, 0); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<3, true>(p, l, b3, L, h, lh * p.d, lh); }, 0, false,
            s0);
      // combined-function bends: the curvature terms of PolyB[1] and PolyB[2],
      // at centimetre amplitudes where the cubic terms carry weight
      const double s1[6] = {U(-2e-2, 2e-2), U(-1e-3, 1e-3), U(-1.5e-2, 1.5e-2), U(-1e-3, 1e-3), s0[4], d};
      const double g1[2] = {0.0, -0.35};
      check("bend kick gradient", [=](Part &p, double) { bend_kick_c<1, false>(p, g1, L, h, lh * p.d, lh, -0.35, 0); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<1, false>(p, l, g1, L, h, lh * p.d, lh, -0.35, 0); }, 0,
            false, s1);
      const double g3[4] = {2e-4, 0.4, -6.0, 25.0};
      check("bend kick grad+sext MO=3",
            [=](Part &p, double) { bend_kick_c<3, false>(p, g3, L, h, lh * p.d, lh, 0.4, -6.0); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<3, false>(p, l, g3, L, h, lh * p.d, lh, 0.4, -6.0); },
            0, false, s1);
      // orders above the record's padded with zeros (the MO = 9 reverse map
      // serves every MaxNPole above 4)
      const double g5[6] = {0.0, -0.2, 3.0, 0.0, 150.0, -900.0};
      double g9[10] = {0.0, -0.2, 3.0, 0.0, 150.0, -900.0, 0.0, 0.0, 0.0, 0.0};
      check("bend kick MO=5 as MO=9",
            [=](Part &p, double) { bend_kick_c<5, false>(p, g5, L, h, lh * p.d, lh, -0.2, 3.0); },
            [=](Part &p, Adj &l, double *) { bend_kick_rev<9, false>(p, l, g9, L, h, lh * p.d, lh, -0.2, 3.0); },
            0, false, s1);
    }
    {  // multipole kicks with knobs on PolyB[2], PolyB[1], PolyA[3]
      const int kinds[3][2] = {{2, 2}, {2, 1}, {3, 3}};
      for (auto &kk : kinds) {
        const int kind = kk[0], j = kk[1];
        char nm[64];
        std::snprintf(nm, sizeof nm, "mpole kick %s[%d]", kind == 2 ? "B" : "A", j);
        check(nm,
              [=](Part &p, double kv) {
                double b[4] = {1e-4, 0.5, -8.0, 30.0}, a[4] = {2e-4, 0.1, 3.0, -10.0};
                (kind == 2 ? b : a)[j] = kv;
                double re, im;
                horner_reg<3>(b, a, p.x, p.y, re, im);
                p.px -= 0.03 * re;
                p.py += 0.03 * im;
              },
              [=](Part &p, Adj &l, double *g) {
                double b[4] = {1e-4, 0.5, -8.0, 30.0}, a[4] = {2e-4, 0.1, 3.0, -10.0};
                (kind == 2 ? b : a)[j] = 7.5;
                mpole_kick_rev<3>(p, l, b, a, 0.03, Knob{0, kind, j}, g);
              },
              7.5, true, s0, 0.1);
      }
    }
    {  // series quads, focusing and defocusing, K knob
      const std::vector<double> C = quad_coefs(8);
      for (double K : {1.2, -1.1}) {
        const double L = 0.3;
        check(K > 0 ? "quad series K>0" : "quad series K<0",
              [=](Part &p, double kv) { quad_series<8>(p, C.data(), L, L * L, kv > 0, std::fabs(kv)); },
              [=](Part &p, Adj &l, double *g) {
                quad_series_rev<8>(p, l, C.data(), L, L * L, K > 0, std::fabs(K), K > 0 ? 1.0 : -1.0,
                                   Knob{0, 1, 0}, g);
              },
              K, true, s0);
      }
    }
    {  // linear sector bends (bend6): focusing, defocusing and zero-gradient vertical planes
      const double L = 1.5, kx = 0.157 / 1.5;
      for (double grd : {0.3, -0.5, 0.0}) {
        char nm[64];
        std::snprintf(nm, sizeof nm, "bend6 K=%+.1f", grd);
        check(nm, [=](Part &p, double) { bend_linear(p, L, kx, grd); },
              [=](Part &p, Adj &l, double *) { bend_linear_rev(p, l, L, kx, grd); }, 0, false, s0);
      }
    }
    {  // trigonometric quads (u = L^2 |K|/(1+delta) near 7.5, beyond the series range), K knob;
       // cosh sqrt u ~ 15 amplifies the differences' round-off: 1e-5 coordinate steps
      for (double K : {30.0, -30.0}) {
        const double L = 0.5;
        check(K > 0 ? "quad trig K>0" : "quad trig K<0",
              [=](Part &p, double kv) { quad(p, L, kv, std::fabs(kv)); },
              [=](Part &p, Adj &l, double *g) {
                quad_trig_rev(p, l, L, L * L, K > 0, std::fabs(K), K > 0 ? 1.0 : -1.0, Knob{0, 1, 0}, g);
              },
              K, true, s0, 1e-7, 2e-7, 1e-5);
      }
    }
    {
      const double w = TWOPI * 562e6 / C0;
      check("cavity thin", [=](Part &p, double) { cavity(p, 0.0, 5e-4, w, 0.1, true); },
            [=](Part &p, Adj &l, double *) { cavity_rev(p, l, 0.0, 5e-4, w, 0.1, true); }, 0, false, s0);
      check("cavity thick", [=](Part &p, double) { cavity(p, 0.4, 5e-4, w, 0.1, true); },
            [=](Part &p, Adj &l, double *) { cavity_rev(p, l, 0.4, 5e-4, w, 0.1, true); }, 0, false, s0, 1e-7,
            2e-7, 1e-5);
      check("corrector", [=](Part &p, double) { corrector(p, 0.2, 1e-4, -2e-4); },
            [=](Part &p, Adj &l, double *) { corrector_rev(p, l, 0.2, 1e-4, -2e-4); }, 0, false, s0);
    }
  }
  // maps that move delta inside the element
  for (int trial = 0; trial < 3; ++trial) {
    const double d = trial == 0 ? -0.03 : trial == 1 ? 0.0 : 0.027;
    const double s0[6] = {U(-3e-3, 3e-3), U(-3e-4, 3e-4), U(-2e-3, 2e-3), U(-2e-4, 2e-4), U(-1e-2, 1e-2), d};
    const double s1[6] = {U(-2e-2, 2e-2), U(-1e-3, 1e-3), U(-1.5e-2, 1.5e-2), U(-1e-3, 1e-3), s0[4], d};
    auto div = [](double v) { return 1.0 / v; };
    {  // radiating bend kicks (classical loss), a 3 GeV ring's dipole: crad L of order 1e-4
      const double h = 0.157 / 1.5, E0 = 3.0, crad = CGAMMA * E0 * E0 * E0 / (TWOPI * 1e9);
      for (double L : {0.15 * KICK1, 0.15 * KICK2}) {
        const double b0[1] = {0.0};
        check(L > 0 ? "rad bend kick MO=0" : "rad bend kick MO=0 (L<0)",
              [=](Part &p, double) { bend_kick_rad(p, b0, 0, L, h, 0.0, 0.0, crad, 0.0, 1, 0.0, div); },
              [=](Part &p, Adj &l, double *) { bend_kick_rad_rev(p, l, b0, 0, L, h, 0.0, 0.0, crad); }, 0, false,
              s0);
        const double g3[4] = {2e-4, 0.4, -6.0, 25.0};
        check("rad bend kick grad+sext",
              [=](Part &p, double) { bend_kick_rad(p, g3, 3, L, h, 0.4, -6.0, crad, 0.0, 1, 0.0, div); },
              [=](Part &p, Adj &l, double *) { bend_kick_rad_rev(p, l, g3, 3, L, h, 0.4, -6.0, crad); }, 0, false,
              s1);
      }
      // a whole radiating element: edges, pseudo-drifts with delta rebuilt per sub-step
      static double P[BM_B + 4];
      const double Lb = 1.5, ns = 3, sl = Lb / ns;
      P[BM_PSI1] = 0.011; P[BM_PSI2] = -0.007; P[BM_IRHO] = h;
      P[BM_L1] = sl * DRIFT1; P[BM_L2] = sl * DRIFT2; P[BM_K1] = sl * KICK1; P[BM_K2] = sl * KICK2;
      P[BM_SL] = sl; P[BM_CRAD] = crad; P[BM_CQ] = 0;
      P[BM_HS1] = h * P[BM_L1]; P[BM_HQ1] = 0.25 * P[BM_HS1] * P[BM_L1];
      P[BM_HS2] = h * P[BM_L2]; P[BM_HQ2] = 0.25 * P[BM_HS2] * P[BM_L2];
      P[BM_K1C] = 0.4; P[BM_K2C] = -6.0;
      P[BM_B] = 0.0; P[BM_B + 1] = 0.4; P[BM_B + 2] = -6.0; P[BM_B + 3] = 25.0;
      const Hdr hd{OP_BENDMP_RAD, 0, 3, 3};
      check("rad BendMPole element",
            [=](Part &p, double) {
              edge(p, P[BM_PSI1]);
              for (int m = 0; m < 3; ++m) {
                hxdrift(p, P[BM_L1], h, P[BM_HS1], P[BM_HQ1]);
                bend_kick_rad(p, P + BM_B, 3, P[BM_K1], h, 0.4, -6.0, crad, 0.0, 1, 0.0, div);
                hxdrift(p, P[BM_L2], h, P[BM_HS2], P[BM_HQ2]);
                bend_kick_rad(p, P + BM_B, 3, P[BM_K2], h, 0.4, -6.0, crad, 0.0, 1, 0.0, div);
                hxdrift(p, P[BM_L2], h, P[BM_HS2], P[BM_HQ2]);
                bend_kick_rad(p, P + BM_B, 3, P[BM_K1], h, 0.4, -6.0, crad, 0.0, 1, 0.0, div);
                hxdrift(p, P[BM_L1], h, P[BM_HS1], P[BM_HQ1]);
              }
              edge(p, P[BM_PSI2]);
            },
            [=](Part &p, Adj &l, double *) { bendmp_rad_rev(p, l, hd, P); }, 0, false, s1);
    }
    {  // lumped damping
      static const double C[7] = {0.4, 1 - 2.1e-4, 1 - 1.6e-4, 1 - 3.3e-4, 0, 0, 0};
      check("syncrad damping", [=](Part &p, double) { syncrad(p, C, true, false, 0, 0, 0); },
            [=](Part &p, Adj &l, double *) { syncrad_rev(p, l, C, true); }, 0, false, s0);
    }
    {  // crab cavity at 1.5 GHz, strong enough that the energy term matters
      static const double C[7] = {0.3, 1.5 / 3000.0 * 50, 0.5 / 3000.0 * 50, TWOPI * 1.5e9 / C0, 0.3, 0, 0};
      static const double T[7] = {0.0, 1.5 / 3000.0 * 50, 0.5 / 3000.0 * 50, TWOPI * 1.5e9 / C0, 0.3, 0, 0};
      check("crab cavity thick", [=](Part &p, double) { crabcav(p, C, 1, 0, 0, 0); },
            [=](Part &p, Adj &l, double *) { crabcav_rev(p, l, C, true); }, 0, false, s0);
      check("crab cavity thin", [=](Part &p, double) { crabcav(p, T, 1, 0, 0, 0); },
            [=](Part &p, Adj &l, double *) { crabcav_rev(p, l, T, true); }, 0, false, s0);
    }
    {  // RF multipole, normal and skew orders 0..3, at centimetre amplitudes
      static const double C[12] = {0.3, TWOPI * 5e8 / C0, 0.2, 0.0, 1e-4, 2e-3, 0.5, 3.0, 0.0, 1e-3, 0.2, -1.5};
      check("rf multipole", [=](Part &p, double) { rfmult(p, C, 4, false, 0.0); },
            [=](Part &p, Adj &l, double *) { rfmult_rev(p, l, C, 4); }, 0, false, s1);
      static const double T[12] = {0.0, TWOPI * 5e8 / C0, -0.4, 0.0, 0.0, 2e-3, 0.5, 3.0, 1e-4, 1e-3, 0.2, -1.5};
      check("rf multipole thin", [=](Part &p, double) { rfmult(p, T, 4, false, 0.0); },
            [=](Part &p, Adj &l, double *) { rfmult_rev(p, l, T, 4); }, 0, false, s0);
    }
    {  // ideal wigglers, both planes, with and without the classical loss
      const double lam = 0.1, k = TWOPI / lam, hw = 0.5, Lw = 2.0, nsl = 20;
      const double crad = CGAMMA * 27.0 / (TWOPI * 1e9);
      static double C[6];
      C[0] = Lw; C[1] = k * k; C[2] = hw * hw; C[3] = Lw / nsl; C[4] = 0.5 * Lw / nsl;
      C[5] = crad * hw * hw * Lw / nsl / 2.0;
      for (int type = 0; type < 2; ++type)
        for (int rad = 0; rad < 2; ++rad) {
          char nm[64];
          std::snprintf(nm, sizeof nm, "ideal wiggler type %d%s", type, rad ? " rad" : "");
          check(nm, [=](Part &p, double) { idealwig(p, C, 20, type, rad != 0); },
                [=](Part &p, Adj &l, double *) { idealwig_rev(p, l, C, 20, type, rad != 0); }, 0, false, s1);
        }
    }
    {  // kick tables, thin and in 4 slices, on a non-uniform grid
      const int mx = 7, my = 5;
      static std::vector<double> C;
      C.assign(2 + mx + my + 2 * mx * my, 0.0);
      C[0] = 0.25;
      C[1] = 4;
      const double xs[mx] = {-0.03, -0.02, -0.008, 0.0, 0.011, 0.02, 0.03};
      const double ys[my] = {-0.02, -0.01, 0.0, 0.009, 0.02};
      for (int i = 0; i < mx; ++i) C[2 + i] = xs[i];
      for (int j = 0; j < my; ++j) C[2 + mx + j] = ys[j];
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          C[2 + mx + my + j * mx + i] = 1e-4 * std::sin(90 * xs[i]) * std::cosh(60 * ys[j]);
          C[2 + mx + my + mx * my + j * mx + i] = 1e-4 * std::cos(70 * xs[i]) * std::sinh(80 * ys[j]);
        }
      check("wig table thin", [=](Part &p, double) { wigtable(p, C.data(), 0, mx, my); },
            [=](Part &p, Adj &l, double *) { wigtable_rev(p, l, C.data(), 0, mx, my); }, 0, false, s0);
      check("wig table sliced", [=](Part &p, double) { wigtable(p, C.data(), 4, mx, my); },
            [=](Part &p, Adj &l, double *) { wigtable_rev(p, l, C.data(), 4, mx, my); }, 0, false, s0);
    }
  }
  for (int trial = 0; trial < 3; ++trial) {
    const double d = trial == 0 ? -0.03 : trial == 1 ? 0.0 : 0.027;
    const double s0[6] = {U(-3e-3, 3e-3), U(-3e-4, 3e-4), U(-2e-3, 2e-3), U(-2e-4, 2e-4), U(-1e-3, 1e-3), d};
    vector_forms(s0);
  }
  std::printf("%s\n", failures ? "FAILED" : "all adjoint primitives ok");
  return failures ? 1 : 0;
}
