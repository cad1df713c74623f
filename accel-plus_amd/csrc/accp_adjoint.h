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
// Reference maps differentiated: fastdrift (atpass.cpp:416-428), quad6
// (:15-72, power-series and trigonometric forms), ATbendhxdrift6 (:343-355),
// bndthinkick (:199-235, normal multipoles with the curvature terms of
// PolyB[1] and PolyB[2]), strthinkick (:358-414), edge
// (:190-196), ACavity::pass (accel.cpp:903-923), ACorrector::pass
// (accel.cpp:270-292).
#pragma once
#include "accp_physics.h"

namespace accp {

struct Adj {
  double x, px, y, py, z, d;
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
ACCP_HD void drift_c_rev(Part &p, Adj &l, double nl, double zc) {
  const double px = p.px, py = p.py, s = px * px + py * py;
  p.x -= nl * px;
  p.y -= nl * py;
  p.z += zc * s;
  const double ipd = p.ipd;
  const double dnl = -nl * ipd, dzc = -2.0 * zc * ipd;
  l.d += (l.x * px + l.y * py) * dnl - l.z * s * dzc;
  l.px += l.x * nl - 2.0 * l.z * zc * px;
  l.py += l.y * nl - 2.0 * l.z * zc * py;
}

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
template <int HI>
ACCP_HD void mpole_kick_rev(Part &p, Adj &l, const double (&b)[HI + 1], const double (&a)[HI + 1],
                            double K, const Knob &kn, double *grad) {
  const double x = p.x, y = p.y;
  double re, im, dre, dim;
  horner_d<HI>(b, a, x, y, re, im, dre, dim);
  p.px += K * re;
  p.py -= K * im;
  const double lpx = l.px, lpy = l.py;
  l.x += K * (-lpx * dre + lpy * dim);
  l.y += K * (lpx * dim + lpy * dre);
  if (kn.kind >= 2) {
    double wr, wi;
    cpow(x, y, kn.j, wr, wi);
    grad[kn.slot] += kn.kind == 2 ? K * (-lpx * wr + lpy * wi) : K * (lpx * wi + lpy * wr);
  }
}

// mpole kick when the element's field is empty but a knob sits on it: the
// map is the identity, only the knob derivative survives.
ACCP_HD void mpole_knob_only(const Part &p, const Adj &l, double K, const Knob &kn, double *grad) {
  if (kn.kind < 2) return;
  double wr, wi;
  cpow(p.x, p.y, kn.j, wr, wi);
  grad[kn.slot] += kn.kind == 2 ? K * (-l.px * wr + l.py * wi) : K * (l.px * wi + l.py * wr);
}

// ---- hard-edge fringe ---------------------------------------------------------
ACCP_HD void edge_rev(Part &p, Adj &l, double psi) {
  p.px -= p.x * psi;
  p.py += p.y * psi;
  l.x += l.px * psi;
  l.y -= l.py * psi;
}

// ---- ATbendhxdrift6 with delta fixed (hxdrift_c) -------------------------------
ACCP_HD void hxdrift_c_rev(Part &p, Adj &l, const HxC &c, double h, double c4) {
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
  // VJP at the input state
  const Adj o = l;
  l.x = o.x * (1.0 + hpc1) + o.y * h * py * c.c1 * f - o.z * h * c.c5 * s;
  l.px = o.x * (a * c.c1 + 2.0 * c.c2 * px) + o.px * (1.0 - 2.0 * c.c3 * px) + o.y * a * py * c.c1 * c4 -
         o.z * 2.0 * a * c.c5 * px;
  l.py = -o.x * 2.0 * c.c2 * py - o.px * 2.0 * c.c3 * py + o.py + o.y * a * c.c1 * f -
         o.z * 2.0 * a * c.c5 * py;
  l.d = o.d + (-o.x * (a * px * c.c1 + 2.0 * c.c2 * d2) + o.px * s * c.c3 - o.y * a * py * f * c.c1 +
               o.z * 2.0 * a * c.c5 * s) * i1;
  p.x = x;
  p.px = px;
  p.y = y;
  p.z = z;
}

// ---- bndthinkick, normal multipoles only, delta fixed (bend_kick_c<MO,PLAIN>) ----
// PLAIN: PolyB[1] = PolyB[2] = 0, so the curvature terms vanish.  Otherwise
// (atpass.cpp:231-233) the px decrement F gains h (k1c (x^2 - y^2/2) +
// k2c (x^3 - 4/3 x y^2)) and the py increment G gains h (k1c x y +
// 4/3 k2c x^2 y + c3 y^3), c3 = h/6 k1c - k2c/3; their x and y derivatives
// join those of the multipole sum (Cauchy-Riemann: d re/dy = -Im P',
// d im/dx = Im P').
template <int MO, bool PLAIN>
ACCP_HD void bend_kick_rev(Part &p, Adj &l, const double (&b)[MO + 1], double L, double h,
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
    const double lpx = l.px, lpy = MO > 0 ? l.py : 0.0;
    l.x += -lpx * L * (dre + h * h) + lpy * L * dim - l.z * lh;
    if (MO > 0) l.y += lpx * L * dim + lpy * L * dre;
    l.d += lpx * lh;
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
    const double lpx = l.px, lpy = l.py;
    l.x += L * (-lpx * fx + lpy * gx) - l.z * lh;
    l.y += L * (-lpx * fy + lpy * gy);
    l.d += lpx * lh;
  }
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
ACCP_HD void quad_rev_core(Part &p, Adj &l, double c, double ch, double s, double sh, double sd, double shd,
                           double L, double L2, bool foc, double g, double sK, const Knob &kn, double *grad) {
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
  const Adj o = l;
  l.x = o.x * mhd + o.px * opd * m21 + o.z * Zx;
  l.px = (o.x * m12 + o.px * opd * mhd + o.z * Zxp) * ipd;
  l.y = o.y * mvd + o.py * opd * m43 + o.z * Zy;
  l.py = (o.y * m34 + o.py * opd * mvd + o.z * Zyp) * ipd;
  l.z = o.z;
  l.d = o.d + o.x * X_d + o.px * PX_d + o.y * Y_d + o.py * PY_d + o.z * Z_d;
  if (kn.kind == 1) {
    const double g_k = sK * ipd;  // dg/dK
    grad[kn.slot] += (o.x * X_g + o.px * PXg + o.y * Y_g + o.py * PYg - o.z * T_g) * g_k;
  }
  p.x = x;
  p.px = xp * opd;
  p.y = y;
  p.py = yp * opd;
  p.z = z;
}

template <int NT, class PT>
ACCP_HD void quad_series_rev(Part &p, Adj &l, const PT *C, double L, double L2, bool foc, double aK,
                             double sK, const Knob &kn, double *grad) {
  const double g = aK * p.ipd;
  const double u = L2 * g;
  double c, ch, s, sh, sd, shd;
  cs_series<NT>(C, u, c, ch, s, sh);
  ds_series<NT>(C, u, sd, shd);
  quad_rev_core(p, l, c, ch, s, sh, sd, shd, L, L2, foc, g, sK, kn, grad);
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
  quad_rev_core(p, l, c, ch, s, sh, (c - s) / (2.0 * u), (ch - sh) / (2.0 * u), L, L2, foc, g, sK, kn, grad);
}

// ---- RF cavity (ACavity::pass) ---------------------------------------------------
ACCP_HD void drift_nl_rev(Part &p, Adj &l, double L) {
  // drift_nl(p, L*ipd): nl = L/(1+d), zc = nl/(2(1+d))
  const double nl = L * p.ipd;
  drift_c_rev(p, l, nl, nl * (0.5 * p.ipd));
}

ACCP_HD void cavity_rev(Part &p, Adj &l, double L, double nv, double w_c0, double phis, bool on) {
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  if (on) {
    const double arg = w_c0 * p.z + phis;
    set_delta(p, p.d - nv * sin(arg));
    l.z += l.d * nv * w_c0 * cos(arg);
  }
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

// ---- ACorrector::pass -------------------------------------------------------------
ACCP_HD void corrector_rev(Part &p, Adj &l, double L, double kx, double ky) {
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
  p.px -= kx * p.ipd;
  p.py -= ky * p.ipd;
  l.d += -(l.px * kx + l.py * ky) * p.ipd * p.ipd;
  if (L != 0) drift_nl_rev(p, l, 0.5 * L);
}

}  // namespace accp
