// accp_adjoint.h -- reverse-mode turn map: inverse maps and vector-Jacobian
// products (VJPs) of the element maps in accp_physics.h.
//
// Each *_rev function takes the state AFTER the element (or sub-step),
// replaces it by the state before it (the maps are symplectic, hence
// invertible in closed form), and replaces the adjoint l = dL/d(state after)
// by dL/d(state before), adding parameter derivatives where the element
// carries a differentiated knob.  Inside an element without radiation delta
// is fixed, so the delta-dependent constants are rebuilt from 1/(1+delta)
// and their delta-derivatives enter l.d.
//
// Reference maps differentiated: fastdrift (atpass.cpp:416-428), drift
// (:6-13, quad6 at K = 0), quad6 (:15-72, power-series and trigonometric
// forms), bend6 (:76-187),
// ATbendhxdrift6 (:343-355),
// bndthinkick (:199-235, normal multipoles with the curvature terms of
// PolyB[1] and PolyB[2]), strthinkick (:358-414), edge
// (:190-196), AMatrix::pass (accel.cpp:1697-1729), ACavity::pass
// (accel.cpp:903-923), ACorrector::pass (accel.cpp:270-292); with delta
// changing inside the element: bndthinkickrad with the classical loss
// (atpass.cpp:255-339, flag 1), ACrabCav::pass and ARFMult::pass without
// noise (accel.cpp:968-1034, 1146-1201), ASyncRad::pass with damping only
// (accel.cpp:1813-1854), AIdealWig::pass with and without its classical
// loss (accel.cpp:1463-1537) and AWigTable::pass (accel.cpp:1338-1382).
// The loss maps keep x and y and move delta by a function of the angles
// and of delta itself; given the angles after the loss, the delta before it
// is the root near 1 + delta of a quadratic, so these inverses stay in
// closed form too.
//
// The maps of the lean sweep (drift, edge, dipole pseudo-drift and kick,
// multipole kick, series quad, K = 0 quad, cavity, corrector) carry NV
// adjoint vectors at once (the *_n<NV> forms): the inverse and the partial
// derivatives are formed once, then applied to each vector; knob
// derivatives go through ga(vector, slot, value).  The plain names are the
// one-vector forms.
#pragma once
#include "accp_physics.h"

namespace accp {

struct Adj {
  double x, px, y, py, z, d;
};

// One-vector adapters: a single adjoint as an array of one, and a knob
// gradient array as an accumulator.
ACCP_HD Adj (&as_vec(Adj &l))[1] { return *reinterpret_cast<Adj(*)[1]>(&l); }
struct GradAcc1 {
  double *g;
  ACCP_HD void operator()(int, int slot, double v) const { g[slot] += v; }
};

// Knob of one element: which gradient slot it feeds and what it moves.
// kind: 0 none, 1 Quad K, 2 PolyB[j], 3 PolyA[j].
struct Knob {
  int slot;
  int kind;
  int j;
};

// ---- drift with delta fixed: drift_c(p, nl, zc), nl = L/(1+d),
// zc = nl/(2(1+d)) ------------------------------------------------------------
template <int NV>
ACCP_HD void drift_c_rev_n(Part &p, Adj (&l)[NV], double nl, double zc) {
  const double px = p.px, py = p.py, s = px * px + py * py;
  p.x -= nl * px;
  p.y -= nl * py;
  p.z += zc * s;
  const double ipd = p.ipd;
  const double dnl = -nl * ipd, zs = -2.0 * zc * ipd * s, zpx = 2.0 * zc * px, zpy = 2.0 * zc * py;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    l[v].d += (l[v].x * px + l[v].y * py) * dnl - l[v].z * zs;
    l[v].px += l[v].x * nl - l[v].z * zpx;
    l[v].py += l[v].y * nl - l[v].z * zpy;
  }
}
ACCP_HD void drift_c_rev(Part &p, Adj &l, double nl, double zc) { drift_c_rev_n<1>(p, as_vec(l), nl, zc); }

// ---- complex polynomial P(w) = sum (b_n + i a_n) w^n and P'(w) ------------
template <int HI>
ACCP_HD void horner_d(const double (&b)[HI + 1], const double (&a)[HI + 1], double x, double y,
                      double &re, double &im, double &dre, double &dim) {
  re = b[HI];
  im = a[HI];
  dre = 0.0;
  dim = 0.0;
#pragma unroll
  for (int i = HI - 1; i >= 0; --i) {
    // derivative first (uses the old P), then P
    const double tdr = dre * x - dim * y + re;
    dim = dim * x + dre * y + im;
    dre = tdr;
    const double t = re * x - im * y + b[i];
    im = im * x + re * y + a[i];
    re = t;
  }
}

// w^j for the knob derivative.
ACCP_HD void cpow(double x, double y, int j, double &wr, double &wi) {
  wr = 1.0;
  wi = 0.0;
  for (int i = 0; i < j; ++i) {
    const double t = wr * x - wi * y;
    wi = wr * y + wi * x;
    wr = t;
  }
}

// strthinkick with kick length K: px -= K Re P, py += K Im P.
template <int NV, int HI, class GA>
ACCP_HD void mpole_kick_rev_n(Part &p, Adj (&l)[NV], const double (&b)[HI + 1], const double (&a)[HI + 1],
                              double K, const Knob &kn, const GA &ga) {
  const double x = p.x, y = p.y;
  double re, im, dre, dim;
  horner_d<HI>(b, a, x, y, re, im, dre, dim);
  p.px += K * re;
  p.py -= K * im;
  const double kr = K * dre, ki = K * dim;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double lpx = l[v].px, lpy = l[v].py;
    l[v].x += -lpx * kr + lpy * ki;
    l[v].y += lpx * ki + lpy * kr;
  }
  if (kn.kind >= 2) {
    double wr, wi;
    cpow(x, y, kn.j, wr, wi);
    // (cx, cy) = d(px, py)/d(knob): (-K wr, K wi) for PolyB[j], (K wi, K wr)
    // for PolyA[j] (d P/d b_j = w^j, d P/d a_j = i w^j)
    const double cx = kn.kind == 2 ? -K * wr : K * wi, cy = kn.kind == 2 ? K * wi : K * wr;
#pragma unroll
    for (int v = 0; v < NV; ++v) ga(v, kn.slot, l[v].px * cx + l[v].py * cy);
  }
}
template <int HI>
ACCP_HD void mpole_kick_rev(Part &p, Adj &l, const double (&b)[HI + 1], const double (&a)[HI + 1],
                            double K, const Knob &kn, double *grad) {
  mpole_kick_rev_n<1, HI>(p, as_vec(l), b, a, K, kn, GradAcc1{grad});
}

// mpole kick when the element's field is empty but a knob sits on it: the
// map is the identity, only the knob derivative survives.
template <int NV, class GA>
ACCP_HD void mpole_knob_only_n(const Part &p, const Adj (&l)[NV], double K, const Knob &kn, const GA &ga) {
  if (kn.kind < 2) return;
  double wr, wi;
  cpow(p.x, p.y, kn.j, wr, wi);
  const double cx = kn.kind == 2 ? -K * wr : K * wi, cy = kn.kind == 2 ? K * wi : K * wr;
#pragma unroll
  for (int v = 0; v < NV; ++v) ga(v, kn.slot, l[v].px * cx + l[v].py * cy);
}
ACCP_HD void mpole_knob_only(const Part &p, const Adj &l, double K, const Knob &kn, double *grad) {
  mpole_knob_only_n<1>(p, *reinterpret_cast<const Adj(*)[1]>(&l), K, kn, GradAcc1{grad});
}

// ---- hard-edge fringe ---------------------------------------------------------
template <int NV>
ACCP_HD void edge_rev_n(Part &p, Adj (&l)[NV], double psi) {
  p.px -= p.x * psi;
  p.py += p.y * psi;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    l[v].x += l[v].px * psi;
    l[v].y -= l[v].py * psi;
  }
}
ACCP_HD void edge_rev(Part &p, Adj &l, double psi) { edge_rev_n<1>(p, as_vec(l), psi); }

// ---- ATbendhxdrift6 with delta fixed (hxdrift_c) -------------------------------
template <int NV>
ACCP_HD void hxdrift_c_rev_n(Part &p, Adj (&l)[NV], const HxC &c, double h, double c4) {
  const double i1 = p.ipd;
  // inverse: py unchanged; px solves px = q + c3 px^2 with q = px' + c3 py^2.
  // The fixed point contracts by 2 c3 |px| (about 1e-5 for a bend slice and
  // mrad angles), so four iterations reach double precision without the
  // square root and division of the closed form.
  const double py = p.py;
  const double q = p.px + c.c3 * py * py;
  double px = q;
#pragma unroll
  for (int it = 0; it < 4; ++it) px = q + c.c3 * px * px;
  const double px2 = px * px, py2 = py * py, s = px2 + py2, d2 = px2 - py2;
  const double hpc1 = h * px * c.c1;
  // x = (x' - px c1 - c2 d2) / (1 + e), e = h px c1 tiny: 1/(1+e) as a series
  const double inv = 1.0 - hpc1 * (1.0 - hpc1 * (1.0 - hpc1 * (1.0 - hpc1)));
  const double x = (p.x - px * c.c1 - c.c2 * d2) * inv;
  const double a = 1.0 + h * x;
  const double f = 1.0 + px * c4;
  const double y = p.y - a * py * c.c1 * f;
  const double z = p.z + a * c.c5 * s;
  // VJP at the input state: the partial derivatives of (x, px, y, z)
  // after the step with respect to (x, px, py, delta) before it
  const double ac1 = a * c.c1, a2c5 = 2.0 * a * c.c5;
  const double xx = 1.0 + hpc1, yx = h * py * c.c1 * f, zx = -h * c.c5 * s;
  const double xpx = ac1 + 2.0 * c.c2 * px, pxpx = 1.0 - 2.0 * c.c3 * px, ypx = ac1 * py * c4, zpx = -a2c5 * px;
  const double xpy = -2.0 * c.c2 * py, pxpy = -2.0 * c.c3 * py, ypy = ac1 * f, zpy = -a2c5 * py;
  const double xd = -(ac1 * px + 2.0 * c.c2 * d2) * i1, pxd = s * c.c3 * i1, yd = -ac1 * py * f * i1,
               zd = a2c5 * s * i1;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const Adj o = l[v];
    l[v].x = o.x * xx + o.y * yx + o.z * zx;
    l[v].px = o.x * xpx + o.px * pxpx + o.y * ypx + o.z * zpx;
    l[v].py = o.x * xpy + o.px * pxpy + o.py + o.y * ypy + o.z * zpy;
    l[v].d = o.d + o.x * xd + o.px * pxd + o.y * yd + o.z * zd;
  }
  p.x = x;
  p.px = px;
  p.y = y;
  p.z = z;
}
ACCP_HD void hxdrift_c_rev(Part &p, Adj &l, const HxC &c, double h, double c4) {
  hxdrift_c_rev_n<1>(p, as_vec(l), c, h, c4);
}

// ---- bndthinkick, normal multipoles only, delta fixed (bend_kick_c<MO,PLAIN>) ----
// PLAIN: PolyB[1] = PolyB[2] = 0, so the curvature terms vanish.  Otherwise
// (atpass.cpp:231-233) the px decrement F gains h (k1c (x^2 - y^2/2) +
// k2c (x^3 - 4/3 x y^2)) and the py increment G gains h (k1c x y +
// 4/3 k2c x^2 y + c3 y^3), c3 = h/6 k1c - k2c/3; their x and y derivatives
// join those of the multipole sum (Cauchy-Riemann: d re/dy = -Im P',
// d im/dx = Im P').
template <int NV, int MO, bool PLAIN>
ACCP_HD void bend_kick_rev_n(Part &p, Adj (&l)[NV], const double (&b)[MO + 1], double L, double h,
                             double lhd, double lh, double k1c = 0.0, double k2c = 0.0) {
  const double x = p.x, y = p.y;
  double a0[MO + 1];
#pragma unroll
  for (int i = 0; i <= MO; ++i) a0[i] = 0.0;
  double re, im, dre, dim;
  horner_d<MO>(b, a0, x, y, re, im, dre, dim);
  if (PLAIN) {
    p.px += L * (re + h * (h * x)) - lhd;
    if (MO > 0) p.py -= L * im;
    p.z += lh * x;
    const double cxx = -L * (dre + h * h), cxy = L * dim, cyy = L * dre;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double lpx = l[v].px, lpy = MO > 0 ? l[v].py : 0.0;
      l[v].x += lpx * cxx + lpy * cxy - l[v].z * lh;
      if (MO > 0) l[v].y += lpx * cxy + lpy * cyy;
      l[v].d += lpx * lh;
    }
  } else {
    const double c3 = h / 6.0 * k1c - k2c / 3.0;
    p.px += L * (re + h * (h * x + k1c * (x * x - 0.5 * y * y) + k2c * (x * x * x - (4.0 / 3.0) * x * y * y))) -
            lhd;
    p.py -= L * (im + h * (k1c * x * y + (4.0 / 3.0) * k2c * x * x * y + c3 * y * y * y));
    p.z += lh * x;
    const double fx = dre + h * (h + 2.0 * k1c * x + k2c * (3.0 * x * x - (4.0 / 3.0) * y * y));
    const double fy = -dim - h * (k1c * y + (8.0 / 3.0) * k2c * x * y);
    const double gx = dim + h * (k1c * y + (8.0 / 3.0) * k2c * x * y);
    const double gy = dre + h * (k1c * x + (4.0 / 3.0) * k2c * x * x + 3.0 * c3 * y * y);
    const double lfx = -L * fx, lgx = L * gx, lfy = -L * fy, lgy = L * gy;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double lpx = l[v].px, lpy = l[v].py;
      l[v].x += lpx * lfx + lpy * lgx - l[v].z * lh;
      l[v].y += lpx * lfy + lpy * lgy;
      l[v].d += lpx * lh;
    }
  }
}
template <int MO, bool PLAIN>
ACCP_HD void bend_kick_rev(Part &p, Adj &l, const double (&b)[MO + 1], double L, double h,
                           double lhd, double lh, double k1c = 0.0, double k2c = 0.0) {
  bend_kick_rev_n<1, MO, PLAIN>(p, as_vec(l), b, L, h, lhd, lh, k1c, k2c);
}

// ---- quad6 from the power series, with the K derivative ----------------------
// D(v) = sum (j+1) v^j/(2j+3)! gives d/du [sin sqrt u / sqrt u] = -D(-u) and
// d/du [sinh sqrt u / sqrt u] = D(u); its coefficients follow the four blocks
// of the forward series in the element record (two more blocks of M).
template <int NT, class PT>
ACCP_HD void ds_series(const PT *C, double u, double &sd, double &shd) {
  constexpr int M = NT / 2;
  const double w = u * u;
  double de = C[5 * M - 1], dd = C[6 * M - 1];
#pragma unroll
  for (int m = M - 2; m >= 0; --m) {
    de = de * w + C[4 * M + m];
    dd = dd * w + C[5 * M + m];
  }
  sd = -(de - u * dd);
  shd = de + u * dd;
}

// quad6 reverse map from the functions of u = L^2 g, g = |K|/(1+delta):
// c = cos sqrt u, ch = cosh sqrt u, s = sin sqrt u / sqrt u,
// sh = sinh sqrt u / sqrt u and sd, shd = d s/du, d sh/du.
template <int NV, class GA>
ACCP_HD void quad_rev_core_n(Part &p, Adj (&l)[NV], double c, double ch, double s, double sh, double sd,
                             double shd, double L, double L2, bool foc, double g, double sK, const Knob &kn,
                             const GA &ga) {
  const double ipd = p.ipd, opd = 1.0 + p.d;
  // matrix and d/dg of its elements
  double mhd, m12, m21, mvd, m34, m43, mhd_g, m12_g, m21_g, mvd_g, m34_g, m43_g;
  const double cg = -0.5 * s * L2, chg = 0.5 * sh * L2, sg = sd * L2, shg = shd * L2;
  if (foc) {
    mhd = c; m12 = L * s; m21 = -m12 * g;
    mvd = ch; m34 = L * sh; m43 = m34 * g;
    mhd_g = cg; m12_g = L * sg; m21_g = -m12_g * g - m12;
    mvd_g = chg; m34_g = L * shg; m43_g = m34_g * g + m34;
  } else {
    mhd = ch; m12 = L * sh; m21 = m12 * g;
    mvd = c; m34 = L * s; m43 = -m34 * g;
    mhd_g = chg; m12_g = L * shg; m21_g = m12_g * g + m12;
    mvd_g = cg; m34_g = L * sg; m43_g = -m34_g * g - m34;
  }
  // inverse of the linear part (det = 1): angles first
  const double XP = p.px * ipd, YP = p.py * ipd;
  const double x = mhd * p.x - m12 * XP, xp = -m21 * p.x + mhd * XP;
  const double y = mvd * p.y - m34 * YP, yp = -m43 * p.y + mvd * YP;
  const double A = L - mhd * m12, B = L - mvd * m34, Cc = L + mhd * m12, D = L + mvd * m34;
  const double T1 = g * (x * x * A - y * y * B) * 0.25;
  const double T2 = (xp * xp * Cc + yp * yp * D) * 0.25;
  const double T3 = (x * xp * m12 * m21 + y * yp * m34 * m43) * 0.5;
  const double z = p.z + T1 + T2 + T3;
  // partial derivatives of the outputs at the input state
  const double A_g = -(mhd_g * m12 + mhd * m12_g), B_g = -(mvd_g * m34 + mvd * m34_g);
  const double T_g = (x * x * A - y * y * B) * 0.25 + g * (x * x * A_g - y * y * B_g) * 0.25 -
                     (xp * xp * A_g + yp * yp * B_g) * 0.25 +
                     (x * xp * (m12_g * m21 + m12 * m21_g) + y * yp * (m34_g * m43 + m34 * m43_g)) * 0.5;
  const double Zx = -(g * 2.0 * x * A * 0.25) - xp * m12 * m21 * 0.5;
  const double Zxp = -(2.0 * xp * Cc * 0.25) - x * m12 * m21 * 0.5;
  const double Zy = (g * 2.0 * y * B * 0.25) - yp * m34 * m43 * 0.5;
  const double Zyp = -(2.0 * yp * D * 0.25) - y * m34 * m43 * 0.5;
  const double g_d = -g * ipd;  // dg/ddelta
  const double xp_d = -xp * ipd, yp_d = -yp * ipd;
  // X = mhd x + m12 xp ; PX = opd (m21 x + mhd xp) ; same for y
  const double X_g = mhd_g * x + m12_g * xp, Y_g = mvd_g * y + m34_g * yp;
  const double PXg = opd * (m21_g * x + mhd_g * xp), PYg = opd * (m43_g * y + mvd_g * yp);
  const double X_d = X_g * g_d + m12 * xp_d;
  const double PX_d = (m21 * x + mhd * xp) + opd * (PXg / opd * g_d + mhd * xp_d);
  const double Y_d = Y_g * g_d + m34 * yp_d;
  const double PY_d = (m43 * y + mvd * yp) + opd * (PYg / opd * g_d + mvd * yp_d);
  const double Z_d = -T_g * g_d + Zxp * xp_d + Zyp * yp_d;
  const double om21 = opd * m21, omhd = opd * mhd, om43 = opd * m43, omvd = opd * mvd;
  const double g_k = sK * ipd;  // dg/dK
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const Adj o = l[v];
    l[v].x = o.x * mhd + o.px * om21 + o.z * Zx;
    l[v].px = (o.x * m12 + o.px * omhd + o.z * Zxp) * ipd;
    l[v].y = o.y * mvd + o.py * om43 + o.z * Zy;
    l[v].py = (o.y * m34 + o.py * omvd + o.z * Zyp) * ipd;
    l[v].z = o.z;
    l[v].d = o.d + o.x * X_d + o.px * PX_d + o.y * Y_d + o.py * PY_d + o.z * Z_d;
    if (kn.kind == 1) ga(v, kn.slot, (o.x * X_g + o.px * PXg + o.y * Y_g + o.py * PYg - o.z * T_g) * g_k);
  }
  p.x = x;
  p.px = xp * opd;
  p.y = y;
  p.py = yp * opd;
  p.z = z;
}

template <int NV, int NT, class PT, class GA>
ACCP_HD void quad_series_rev_n(Part &p, Adj (&l)[NV], const PT *C, double L, double L2, bool foc, double aK,
                               double sK, const Knob &kn, const GA &ga) {
  const double g = aK * p.ipd;
  const double u = L2 * g;
  double c, ch, s, sh, sd, shd;
  cs_series<NT>(C, u, c, ch, s, sh);
  ds_series<NT>(C, u, sd, shd);
  quad_rev_core_n<NV>(p, l, c, ch, s, sh, sd, shd, L, L2, foc, g, sK, kn, ga);
}
template <int NT, class PT>
ACCP_HD void quad_series_rev(Part &p, Adj &l, const PT *C, double L, double L2, bool foc, double aK,
                             double sK, const Knob &kn, double *grad) {
  quad_series_rev_n<1, NT>(p, as_vec(l), C, L, L2, foc, aK, sK, kn, GradAcc1{grad});
}

// The trigonometric path (quad, accp_physics.h), for u beyond the series
// range: there u >= 6, so d/du (sin sqrt u / sqrt u) = (c - s)/(2u) and
// d/du (sinh sqrt u / sqrt u) = (ch - sh)/(2u) lose nothing to cancellation.
ACCP_HD void quad_trig_rev(Part &p, Adj &l, double L, double L2, bool foc, double aK, double sK, const Knob &kn,
                           double *grad) {
  const double g = aK * p.ipd;
  const double u = L2 * g;
  const double lt = sqrt(u);
  double sn, cs;
  sincos(lt, &sn, &cs);
  const double c = cs, ch = cosh(lt), s = sn / lt, sh = sinh(lt) / lt;
  quad_rev_core_n<1>(p, as_vec(l), c, ch, s, sh, (c - s) / (2.0 * u), (ch - sh) / (2.0 * u), L, L2, foc, g, sK,
                     kn, GradAcc1{grad});
}


// ---- quad6 with K == 0: the reference's drift() (atpass.cpp:6-13) --------------
// x += px L, y += py L, then z -= pn^2 L (px + y^2)^2 / 2 with the updated y
// (its SQR(r6[1]+SQR(r6[2])) quirk), delta fixed.
template <int NV>
ACCP_HD void drift_kzero_rev_n(Part &p, Adj (&l)[NV], double L) {
  const double pn = p.ipd, px = p.px, py = p.py, y1 = p.y;
  const double w = px + y1 * y1, c = pn * pn * L * 0.5;
  p.x -= px * L;
  p.y -= py * L;
  p.z += c * (w * w);
  const double dzw = -2.0 * c * w;  // dz'/dw
  const double zy = dzw * 2.0 * y1, zpy = zy * L, zd = L * pn * pn * pn * (w * w);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const Adj o = l[v];
    l[v].px = o.x * L + o.px + o.z * dzw;
    l[v].y = o.y + o.z * zy;
    l[v].py = o.y * L + o.py + o.z * zpy;
    l[v].d = o.d + o.z * zd;
  }
}
ACCP_HD void drift_kzero_rev(Part &p, Adj &l, double L) { drift_kzero_rev_n<1>(p, as_vec(l), L); }

// ---- bend6: linear sector bend with gradient (ABend, atpass.cpp:76-187) ----
// Its map mixes delta into every output through g1, g2 and the dispersion
// terms, so the vector-Jacobian product is formed from the Jacobian itself:
// six forward evaluations in dual numbers (value and one directional
// derivative), one per input coordinate.
struct Dual {
  double v, d;
};
ACCP_HD Dual dl(double v) { return Dual{v, 0.0}; }
ACCP_HD Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
ACCP_HD Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
ACCP_HD Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
ACCP_HD Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
ACCP_HD Dual operator/(Dual a, Dual b) { return Dual{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
ACCP_HD Dual operator+(Dual a, double b) { return Dual{a.v + b, a.d}; }
ACCP_HD Dual operator+(double a, Dual b) { return Dual{a + b.v, b.d}; }
ACCP_HD Dual operator-(Dual a, double b) { return Dual{a.v - b, a.d}; }
ACCP_HD Dual operator-(double a, Dual b) { return Dual{a - b.v, -b.d}; }
ACCP_HD Dual operator*(Dual a, double b) { return Dual{a.v * b, a.d * b}; }
ACCP_HD Dual operator*(double a, Dual b) { return Dual{a * b.v, a * b.d}; }
ACCP_HD Dual operator/(Dual a, double b) { return Dual{a.v / b, a.d / b}; }
ACCP_HD Dual operator/(double a, Dual b) { return Dual{a / b.v, -a * b.d / (b.v * b.v)}; }
ACCP_HD Dual dsqrt(Dual a) {
  const double r = sqrt(a.v);
  return Dual{r, a.d * 0.5 / r};
}
ACCP_HD void dsincos(Dual a, Dual &s, Dual &c) {
  double sv, cv;
  sincos(a.v, &sv, &cv);
  s = Dual{sv, cv * a.d};
  c = Dual{cv, -sv * a.d};
}
ACCP_HD void dsinhcosh(Dual a, Dual &s, Dual &c) {
  const double sv = sinh(a.v), cv = cosh(a.v);
  s = Dual{sv, cv * a.d};
  c = Dual{cv, sv * a.d};
}

// bend6's outputs from its inputs {x, px, y, py, z, delta} (the operations of
// bend_linear in accp_physics.h); the branch on the focusing sign reads the
// value part, which delta's positive 1/(1+delta) cannot flip.
ACCP_HD void bend6_dual(const Dual (&in)[6], double L, double kx, double grd, Dual (&out)[6]) {
  const Dual pn = 1.0 / (1.0 + in[5]);
  const Dual g1 = (kx * kx + grd) * pn, g2 = -grd * pn;
  Dual mhd, m12, m21, mvd, m34, m43, s1 = dl(0.0), c1 = dl(1.0), sg1 = dl(0.0);
  if (g1.v == 0) {
    mhd = dl(1.0); m12 = dl(L); m21 = dl(0.0);
  } else if (g1.v > 0) {
    sg1 = dsqrt(g1);
    dsincos(L * sg1, s1, c1);
    mhd = c1; m12 = s1 / sg1; m21 = -(s1 * sg1);
  } else {
    sg1 = dsqrt(-g1);
    dsinhcosh(L * sg1, s1, c1);
    mhd = c1; m12 = s1 / sg1; m21 = s1 * sg1;
  }
  if (g2.v == 0) {
    mvd = dl(1.0); m34 = dl(L); m43 = dl(0.0);
  } else if (g2.v > 0) {
    const Dual sg2 = dsqrt(g2);
    Dual s2, c2;
    dsincos(L * sg2, s2, c2);
    mvd = c2; m34 = s2 / sg2; m43 = -(s2 * sg2);
  } else {
    const Dual sg2 = dsqrt(-g2);
    Dual s2, c2;
    dsinhcosh(L * sg2, s2, c2);
    mvd = c2; m34 = s2 / sg2; m43 = s2 * sg2;
  }
  const Dual x = in[0], xp = in[1] * pn, y = in[2], yp = in[3] * pn;
  const Dual opd = 1.0 + in[5];
  const Dual dp = in[5] * pn;
  Dual nx = mhd * x + m12 * xp;
  Dual npx = (m21 * x + mhd * xp) * opd;
  if (g1.v == 0) {
    nx = nx + dp * L * L * kx * 0.5;
    npx = npx + dp * L * kx * opd;
  } else {
    nx = nx + dp * (1.0 - c1) * kx / g1;
    npx = npx + dp * s1 * kx / (sg1 * pn);
  }
  Dual z = in[4] - xp * xp * (L + mhd * m12) * 0.25;
  if (g1.v != 0) {
    z = z - (L - mhd * m12) * (x * x * g1 + dp * dp * kx * kx / g1 - 2.0 * x * kx * dp) * 0.25;
    z = z - m12 * m21 * (x * xp - xp * dp * kx / g1) * 0.5;
    z = z - (kx * x * m12 + xp * (1.0 - mhd) * kx / g1 + dp * (L - m12) * kx * kx / g1);
  }
  z = z - ((L - mvd * m34) * y * y * g2 + yp * yp * (L + mvd * m34)) * 0.25;
  z = z - m34 * m43 * y * yp * 0.5;
  out[0] = nx;
  out[1] = npx;
  out[2] = mvd * y + m34 * yp;
  out[3] = (m43 * y + mvd * yp) * opd;
  out[4] = z;
  out[5] = in[5];
}

// Inverse: delta passes unchanged and the transverse part is affine with a
// unit-determinant matrix, so (x, x') = M^-1 ((x, x')_out - D); z takes back
// the path-length change of the recovered input.  Then l <- J^T l.
ACCP_HD void bend_linear_rev(Part &p, Adj &l, double L, double kx, double grd) {
  const double d = p.d, pn = p.ipd, opd = 1.0 + d;
  Dual in[6], out[6];
  // the matrix and dispersion at this delta: the forward map of (0, 0, 0, 0, 0, d)
  // gives D; unit vectors give M's columns
  for (int k = 0; k < 6; ++k) in[k] = dl(0.0);
  in[5] = dl(d);
  bend6_dual(in, L, kx, grd, out);
  const double Dx = out[0].v, Dxp = out[1].v * pn;
  in[0] = dl(1.0);
  bend6_dual(in, L, kx, grd, out);
  const double mhd = out[0].v - Dx, m21 = out[1].v * pn - Dxp;
  in[0] = dl(0.0);
  in[1] = dl(opd);  // x' = 1
  bend6_dual(in, L, kx, grd, out);
  const double m12 = out[0].v - Dx;
  in[1] = dl(0.0);
  in[2] = dl(1.0);
  bend6_dual(in, L, kx, grd, out);
  const double mvd = out[2].v, m43 = out[3].v * pn;
  in[2] = dl(0.0);
  in[3] = dl(opd);
  bend6_dual(in, L, kx, grd, out);
  const double m34 = out[2].v;
  const double a = p.x - Dx, b = p.px * pn - Dxp;
  const double x = mhd * a - m12 * b, xp = -m21 * a + mhd * b;
  const double y = mvd * p.y - m34 * p.py * pn, yp = -m43 * p.y + mvd * p.py * pn;
  const double v[6] = {x, xp * opd, y, yp * opd, 0.0, d};
  for (int k = 0; k < 6; ++k) in[k] = dl(v[k]);
  bend6_dual(in, L, kx, grd, out);
  const double z = p.z - out[4].v;  // out[4] is the change of z from 0
  // J^T l, one column per input coordinate
  const double lo[6] = {l.x, l.px, l.y, l.py, l.z, l.d};
  double li[6];
  const double vin[6] = {x, xp * opd, y, yp * opd, z, d};
  for (int j = 0; j < 6; ++j) {
    for (int k = 0; k < 6; ++k) in[k] = Dual{vin[k], k == j ? 1.0 : 0.0};
    bend6_dual(in, L, kx, grd, out);
    double acc = 0.0;
    for (int k = 0; k < 6; ++k) acc += lo[k] * out[k].d;
    li[j] = acc;
  }
  l = Adj{li[0], li[1], li[2], li[3], li[4], li[5]};
  p.x = x;
  p.px = xp * opd;
  p.y = y;
  p.py = yp * opd;
  p.z = z;
}

// ---- RF cavity (ACavity::pass) ---------------------------------------------------
template <int NV>
ACCP_HD void drift_nl_rev_n(Part &p, Adj (&l)[NV], double L) {
  // drift_nl(p, L*ipd): nl = L/(1+d), zc = nl/(2(1+d))
  const double nl = L * p.ipd;
  drift_c_rev_n<NV>(p, l, nl, nl * (0.5 * p.ipd));
}
ACCP_HD void drift_nl_rev(Part &p, Adj &l, double L) { drift_nl_rev_n<1>(p, as_vec(l), L); }

// ---- AMatrix: 6x6 map between half drifts (accel.cpp:1697-1729) -------------
// Inverse by Gaussian elimination with partial pivoting on M r_in = r_out;
// the VJP of the matrix is l <- M^T l.  The host rejects a singular or
// nearly singular Matrix before any reverse sweep (smallest pivot below
// 1e-12 of the largest entry).
template <class PT>
ACCP_HD void matrix_rev(Part &p, Adj &l, double L, const PT *m) {
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  double a[6][7];
  const double rhs[6] = {p.x, p.px, p.y, p.py, p.z, p.d};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int j = 0; j < 6; ++j) a[i][j] = m[i * 6 + j];
    a[i][6] = rhs[i];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    int piv = c;
#pragma unroll
    for (int i = c + 1; i < 6; ++i)
      if (fabs(a[i][c]) > fabs(a[piv][c])) piv = i;
#pragma unroll
    for (int j = 0; j < 7; ++j) {  // swap rows c and piv (selects keep indexing static)
      double t = a[c][j];
#pragma unroll
      for (int i = c + 1; i < 6; ++i)
        if (i == piv) {
          a[c][j] = a[i][j];
          a[i][j] = t;
        }
    }
#pragma unroll
    for (int i = c + 1; i < 6; ++i) {
      const double f = a[i][c] / a[c][c];
#pragma unroll
      for (int j = c; j < 7; ++j) a[i][j] -= f * a[c][j];
    }
  }
  double x[6];
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double acc = a[i][6];
#pragma unroll
    for (int j = i + 1; j < 6; ++j) acc -= a[i][j] * x[j];
    x[i] = acc / a[i][i];
  }
  const double lo[6] = {l.x, l.px, l.y, l.py, l.z, l.d};
  double li[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) acc += m[i * 6 + j] * lo[i];
    li[j] = acc;
  }
  l = Adj{li[0], li[1], li[2], li[3], li[4], li[5]};
  p.x = x[0];
  p.px = x[1];
  p.y = x[2];
  p.py = x[3];
  p.z = x[4];
  set_delta(p, x[5]);
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

template <int NV>
ACCP_HD void cavity_rev_n(Part &p, Adj (&l)[NV], double L, double nv, double w_c0, double phis, bool on) {
  if (L != 0) drift_nl_rev_n<NV>(p, l, 0.5 * L);
  if (on) {
    const double arg = w_c0 * p.z + phis;
    set_delta(p, p.d - nv * sin(arg));
    const double dz = nv * w_c0 * cos(arg);
#pragma unroll
    for (int v = 0; v < NV; ++v) l[v].z += l[v].d * dz;
  }
  if (L != 0) drift_nl_rev_n<NV>(p, l, 0.5 * L);
}
ACCP_HD void cavity_rev(Part &p, Adj &l, double L, double nv, double w_c0, double phis, bool on) {
  cavity_rev_n<1>(p, as_vec(l), L, nv, w_c0, phis, on);
}

// ---- ACorrector::pass -------------------------------------------------------------
template <int NV>
ACCP_HD void corrector_rev_n(Part &p, Adj (&l)[NV], double L, double kx, double ky) {
  if (L != 0) drift_nl_rev_n<NV>(p, l, 0.5 * L);
  p.px -= kx * p.ipd;
  p.py -= ky * p.ipd;
  const double i2 = p.ipd * p.ipd;
#pragma unroll
  for (int v = 0; v < NV; ++v) l[v].d += -(l[v].px * kx + l[v].py * ky) * i2;
  if (L != 0) drift_nl_rev_n<NV>(p, l, 0.5 * L);
}
ACCP_HD void corrector_rev(Part &p, Adj &l, double L, double kx, double ky) {
  corrector_rev_n<1>(p, as_vec(l), L, kx, ky);
}

// ---- bndthinkickrad, classical loss (bend_kick_rad with flag 1) ----------------
// Forward: x' = px/(1+d), y' = py/(1+d); d1 = d - C (1+d)^2 with
// C = crad L B2perp(x, x', y, y') (1 + h x + (x'^2 + y'^2)/2); then
// px = x' (1+d1) - L (-h d1 + F(x, y)), py = y' (1+d1) + L G(x, y),
// z -= L h x.  Inverse: x' and y' from the output momenta and d1, C from
// them, and 1+d = 2 (1+d1) / (1 + sqrt(1 - 4 C (1+d1))).  The VJP runs the
// forward operations backwards.
template <class PT>
ACCP_HD void bend_kick_rad_rev(Part &p, Adj &l, const PT *B, int mo, double L, double h, double k1c, double k2c,
                               double crad) {
  const double x = p.x, y = p.y;
  double re = B[mo], im = 0.0, dre = 0.0, dim = 0.0;
  for (int i = mo - 1; i >= 0; --i) {
    const double tdr = dre * x - dim * y + re;
    dim = dim * x + dre * y + im;
    dre = tdr;
    const double t = re * x - im * y + B[i];
    im = im * x + re * y;
    re = t;
  }
  const double hk = h * (k1c * h - k2c);
  im += hk * y * y * y * (1.0 / 6.0);
  re += -k1c * h * y * y * 0.5 + hk * x * y * y * 0.5;
  // partial derivatives of the corrected sums (Cauchy-Riemann for P)
  const double re_x = dre + hk * y * y * 0.5, re_y = -dim - k1c * h * y + hk * x * y;
  const double im_x = dim, im_y = dre + hk * y * y * 0.5;
  const double c3 = h / 6.0 * k1c - k2c / 3.0;
  const double fx = re + h * (h * x + k1c * (x * x - 0.5 * y * y) + k2c * (x * x * x - (4.0 / 3.0) * x * y * y));
  const double gy = im + h * (k1c * x * y + (4.0 / 3.0) * k2c * x * x * y + c3 * y * y * y);
  const double fx_x = re_x + h * (h + 2.0 * k1c * x + k2c * (3.0 * x * x - (4.0 / 3.0) * y * y));
  const double fx_y = re_y - h * (k1c * y + (8.0 / 3.0) * k2c * x * y);
  const double gy_x = im_x + h * (k1c * y + (8.0 / 3.0) * k2c * x * y);
  const double gy_y = im_y + h * (k1c * x + (4.0 / 3.0) * k2c * x * x + 3.0 * c3 * y * y);
  // the state before the kick
  const double d1 = p.d, u1 = 1.0 + d1;
  const double xp = (p.px + L * (-h * d1 + fx)) / u1;
  const double yp = (p.py - L * gy) / u1;
  const double a = 1.0 + x * h;
  const double vq = a * a + xp * xp + yp * yp;
  const double by = re + h, bx = im;
  const double t1 = by * a, t2 = bx * a, t3 = bx * yp - by * xp;
  const double nn = t1 * t1 + t2 * t2 + t3 * t3;
  const double b2p = nn / vq;
  const double S = a + (xp * xp + yp * yp) * 0.5;
  const double C = crad * b2p * S * L;
  const double u0 = 2.0 * u1 / (1.0 + sqrt(1.0 - 4.0 * C * u1));
  // VJP
  const Adj o = l;
  const double gd1 = o.d + o.px * (xp + L * h) + o.py * yp;
  double gxp = o.px * u1, gyp = o.py * u1;
  double gx = -o.px * L * fx_x + o.py * L * gy_x - o.z * L * h;
  double gyy = -o.px * L * fx_y + o.py * L * gy_y;
  // d1 = d - mean, mean = crad L u0^2 b2p S
  const double gmean = -gd1;
  double gu = gmean * crad * L * 2.0 * u0 * b2p * S;
  const double gb2p = gmean * crad * L * u0 * u0 * S;
  const double gS = gmean * crad * L * u0 * u0 * b2p;
  double ga = gS;
  gxp += gS * xp;
  gyp += gS * yp;
  const double gnn = gb2p / vq, gvq = -gb2p * b2p / vq;
  ga += gvq * 2.0 * a;
  gxp += gvq * 2.0 * xp;
  gyp += gvq * 2.0 * yp;
  const double gt1 = 2.0 * t1 * gnn, gt2 = 2.0 * t2 * gnn, gt3 = 2.0 * t3 * gnn;
  const double gby = gt1 * a - gt3 * xp, gbx = gt2 * a + gt3 * yp;
  ga += gt1 * by + gt2 * bx;
  gxp -= gt3 * by;
  gyp += gt3 * bx;
  gx += ga * h + gby * re_x + gbx * im_x;
  gyy += gby * re_y + gbx * im_y;
  const double iu0 = 1.0 / u0;
  gu -= (gxp * xp + gyp * yp) * iu0;
  l.x = o.x + gx;
  l.px = gxp * iu0;
  l.y = o.y + gyy;
  l.py = gyp * iu0;
  l.z = o.z;
  l.d = gd1 + gu;
  p.px = xp * u0;
  p.py = yp * u0;
  p.z += L * h * x;
  set_delta(p, u0 - 1.0);
}

// ABendMPole::pass with the classical loss (flag 1): the 4th-order slices
// walked back, the pseudo-drift constants rebuilt from the delta of each
// sub-step.
template <class PT>
ACCP_HD void bendmp_rad_rev(Part &p, Adj &l, const Hdr h, const PT *P) {
  const int ns = h.a, mo = h.b & 0xff;
  const double psi1 = P[BM_PSI1], psi2 = P[BM_PSI2], h0 = P[BM_IRHO];
  const double l1 = P[BM_L1], l2 = P[BM_L2], k1 = P[BM_K1], k2 = P[BM_K2];
  const double hs1 = P[BM_HS1], hq1 = P[BM_HQ1], hs2 = P[BM_HS2], hq2 = P[BM_HQ2];
  const double k1c = P[BM_K1C], k2c = P[BM_K2C], crad = P[BM_CRAD];
  const PT *B = P + BM_B;
  if (psi2 != 0) edge_rev(p, l, psi2);
  for (int m = 0; m < ns; ++m) {
    hxdrift_c_rev(p, l, hx_consts(p.ipd, l1, hs1, hq1), h0, hs1 * 0.5);
    bend_kick_rad_rev(p, l, B, mo, k1, h0, k1c, k2c, crad);
    hxdrift_c_rev(p, l, hx_consts(p.ipd, l2, hs2, hq2), h0, hs2 * 0.5);
    bend_kick_rad_rev(p, l, B, mo, k2, h0, k1c, k2c, crad);
    hxdrift_c_rev(p, l, hx_consts(p.ipd, l2, hs2, hq2), h0, hs2 * 0.5);
    bend_kick_rad_rev(p, l, B, mo, k1, h0, k1c, k2c, crad);
    hxdrift_c_rev(p, l, hx_consts(p.ipd, l1, hs1, hq1), h0, hs1 * 0.5);
  }
  if (psi1 != 0) edge_rev(p, l, psi1);
}

// ---- ASyncRad, damping only (syncrad with damp, no excitation) -------------
template <class PT>
ACCP_HD void syncrad_rev(Part &p, Adj &l, const PT *C, bool damp) {
  const double L = C[0];
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  if (damp) {
    p.x /= C[1];
    p.y /= C[2];
    set_delta(p, p.d / C[3]);
    l.x *= C[1];
    l.y *= C[2];
    l.d *= C[3];
  }
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

// ---- ACrabCav without noise ---------------------------------------------------
// Kick at om = -k z + phis: px += nvx sin om, py += nvy sin om,
// d1 = d - A/(1+d) with A = (nvy y + nvx x) k cos om; 1+d is the positive
// root of u^2 - (1+d1) u - A = 0.
template <class PT>
ACCP_HD void crabcav_rev(Part &p, Adj &l, const PT *C, bool on) {
  const double L = C[0];
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  if (on) {
    const double nvx = C[1], nvy = C[2], k = C[3];
    double so, co;
    sincos(-k * p.z + C[4], &so, &co);
    const double sxy = nvy * p.y + nvx * p.x;
    const double A = sxy * k * co;
    const double hh = 0.5 * (1.0 + p.d);
    const double u0 = hh + sqrt(hh * hh + A);
    const double iu = 1.0 / u0;
    const Adj o = l;
    l.x = o.x - o.d * nvx * k * co * iu;
    l.y = o.y - o.d * nvy * k * co * iu;
    l.z = o.z - k * co * (o.px * nvx + o.py * nvy) - o.d * sxy * k * k * so * iu;
    l.d = o.d * (1.0 + A * iu * iu);
    p.px -= nvx * so;
    p.py -= nvy * so;
    set_delta(p, u0 - 1.0);
  }
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

// ---- ARFMult without noise -----------------------------------------------------
// In Cartesian form the kick is px += cos om Re G(w), py -= cos om Im G(w),
// d1 = d + k sin om Re F(w)/(1+d), with w = x + iy,
// G(w) = sum_m (B_m - i A_m) w^m and F' = G, F(0) = 0 (the polar sums of
// the reference are these series term by term); 1+d is the root near
// 1+d1 of u^2 - (1+d1) u + k sin om Re F = 0.
template <class PT>
ACCP_HD void rfmult_rev(Part &p, Adj &l, const PT *C, int n) {
  const double L = C[0], k = C[1];
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  const PT *B = C + 4, *A = C + 4 + n;
  const double x = p.x, y = p.y;
  double gr = 0, gi = 0, dr = 0, di = 0, hr = 0, hi = 0;
  for (int m = n - 1; m >= 0; --m) {
    const double tdr = dr * x - di * y + gr;
    di = dr * y + di * x + gi;
    dr = tdr;
    const double tgr = gr * x - gi * y + B[m];
    gi = gr * y + gi * x - A[m];
    gr = tgr;
    const double inv = 1.0 / (m + 1);
    const double thr = hr * x - hi * y + B[m] * inv;
    hi = hr * y + hi * x - A[m] * inv;
    hr = thr;
  }
  const double fr = hr * x - hi * y;  // Re F, F = w H
  double so, co;
  sincos(-k * p.z + C[2], &so, &co);
  const double bq = k * fr * so;
  const double hh = 0.5 * (1.0 + p.d);
  const double u0 = hh + sqrt(hh * hh - bq);
  const double iu = 1.0 / u0;
  const Adj o = l;
  l.x = o.x + co * (o.px * dr - o.py * di) + o.d * k * so * gr * iu;
  l.y = o.y - co * (o.px * di + o.py * dr) - o.d * k * so * gi * iu;
  l.z = o.z + k * so * (o.px * gr - o.py * gi) - o.d * k * k * co * fr * iu;
  l.d = o.d * (1.0 - bq * iu * iu);
  p.px -= co * gr;
  p.py += co * gi;
  set_delta(p, u0 - 1.0);
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

// ---- AIdealWig -------------------------------------------------------------------
// The loss step keeps the angles: d1 = d - rq (1+d)^2, p1 = p (1+d1)/(1+d).
ACCP_HD void wig_loss_rev(Part &p, Adj &l, double rq) {
  const double u1 = 1.0 + p.d;
  const double u0 = 2.0 * u1 / (1.0 + sqrt(1.0 - 4.0 * rq * u1));
  const double r = u1 / u0, dd = 1.0 - 2.0 * rq * u0;
  const double xp = p.px / u1, yp = p.py / u1;
  const Adj o = l;
  l.px = o.px * r;
  l.py = o.py * r;
  l.d = o.d * dd + (o.px * xp + o.py * yp) * (dd - r);
  p.px = xp * u0;
  p.py = yp * u0;
  set_delta(p, u0 - 1.0);
}

template <class PT>
ACCP_HD void idealwig_rev(Part &p, Adj &l, const PT *C, int nslice, int type, bool rad) {
  const double k2 = C[1], hh = C[2], lns = C[3], half = C[4], rq = C[5];
  for (int j = 0; j < nslice; ++j) {
    drift_nl_rev(p, l, half);
    if (rad) wig_loss_rev(p, l, rq);
    if (type == 0) {
      const double y = p.y, y2 = y * y;
      p.py += hh * (0.5 * y + 1.0 / 3.0 * y2 * y * k2 + 1.0 / 15.0 * y2 * y2 * y * k2 * k2) * lns;
      l.y -= l.py * hh * (0.5 + y2 * k2 + 1.0 / 3.0 * y2 * y2 * k2 * k2) * lns;
    } else if (type == 1) {
      const double x = p.x, x2 = x * x;
      p.px += hh * (0.5 * x + 1.0 / 3.0 * x2 * x * k2 + 1.0 / 15.0 * x2 * x2 * x * k2 * k2) * lns;
      l.x -= l.px * hh * (0.5 + x2 * k2 + 1.0 / 3.0 * x2 * x2 * k2 * k2) * lns;
    }
    if (rad) wig_loss_rev(p, l, rq);
    drift_nl_rev(p, l, half);
  }
}

// ---- AWigTable: the bilinear kick and its piecewise-constant gradient ---------
template <class PT>
ACCP_HD void wt_kick_rev(Part &p, Adj &l, const PT *xl, int mx, const PT *yl, int my, const PT *xk,
                         const PT *yk, double w) {
  const double x = p.x, y = p.y;
  const int i = wt_cell(xl, mx, x), j = wt_cell(yl, my, y);
  const double dx = xl[i + 1] - xl[i], dy = yl[j + 1] - yl[j];
  const double t = (x - xl[i]) / dx, u = (y - yl[j]) / dy;
  const int o0 = j * mx + i, o1 = o0 + mx;
  const double ax = ((1. - u) * (xk[o0 + 1] - xk[o0]) + u * (xk[o1 + 1] - xk[o1])) / dx;
  const double ay = ((1. - t) * (xk[o1] - xk[o0]) + t * (xk[o1 + 1] - xk[o0 + 1])) / dy;
  const double bx = ((1. - u) * (yk[o0 + 1] - yk[o0]) + u * (yk[o1 + 1] - yk[o1])) / dx;
  const double by = ((1. - t) * (yk[o1] - yk[o0]) + t * (yk[o1 + 1] - yk[o0 + 1])) / dy;
  p.px -= wt_interp(xl, mx, yl, my, xk, x, y) * w;
  p.py -= wt_interp(xl, mx, yl, my, yk, x, y) * w;
  l.x += w * (l.px * ax + l.py * bx);
  l.y += w * (l.px * ay + l.py * by);
}

template <class PT>
ACCP_HD void wigtable_rev(Part &p, Adj &l, const PT *C, int nslice, int mx, int my) {
  const PT *xl = C + 2, *yl = xl + mx, *xk = yl + my, *yk = xk + mx * my;
  if (nslice == 0) {
    wt_kick_rev(p, l, xl, mx, yl, my, xk, yk, 1.0);
    return;
  }
  const double ns = C[1], hl = C[0] / ns;
  for (int s = 0; s < nslice; ++s) {
    drift_nl_rev(p, l, hl);
    wt_kick_rev(p, l, xl, mx, yl, my, xk, yk, 1.0 / ns);
    drift_nl_rev(p, l, hl);
  }
}

}  // namespace accp
