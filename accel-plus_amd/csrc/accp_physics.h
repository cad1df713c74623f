// accp_physics.h -- element maps of the tracking path, one particle at a time.
//
// Each function restates one reference integrator primitive
// (src/lib/accel/atpass.cpp) on a particle held in registers.  Differences
// from the reference are confined to rounding: 1/(1+delta) is carried with
// the particle and refreshed whenever delta changes (the reference divides
// by (1+delta) in every drift), divisions by constants become products with
// host-computed reciprocals, and hipcc may contract a*b+c into FMA.  The
// per-coordinate tolerance of the parity tests (1e-12 relative after one
// turn) covers these ulp-level effects.
#pragma once
#if !defined(__HIPCC_RTC__)  // runtime compilation provides the device math itself
#include <math.h>
#endif
#include <stdint.h>

#include "accp_table.h"

#if defined(__HIPCC__)
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif
#define ACCP_HD __host__ __device__ __forceinline__
#else
#define ACCP_HD inline
#endif

namespace accp {

struct Part {
  double x, px, y, py, z, d;
  double ipd;  // 1/(1+d)
};

ACCP_HD void set_delta(Part &p, double d) {
  p.d = d;
  p.ipd = 1.0 / (1.0 + d);
}

ACCP_HD bool is_live(const Part &p) { return p.x == p.x; }

// fastdrift with NormL = L/(1+delta) (atpass.cpp:416-428).
ACCP_HD void drift_nl(Part &p, double nl) {
  const double s = p.px * p.px + p.py * p.py;
  p.x += nl * p.px;
  p.y += nl * p.py;
  p.z -= nl * s * (0.5 * p.ipd);
}

// drift() as quad6 uses it for K == 0 (atpass.cpp:6-13, :25-28), including
// the reference's (px + y^2)^2 in the path-length term.
ACCP_HD void drift_kzero(Part &p, double L) {
  const double pn = p.ipd;
  p.x += p.px * L;
  p.y += p.py * L;
  const double w = p.px + p.y * p.y;
  p.z -= pn * pn * L * (w * w) * 0.5;
}

// quad6 (atpass.cpp:15-72) for K != 0; aK = |K|.
ACCP_HD void quad(Part &p, double L, double K, double aK) {
  const double pn = p.ipd;
  const double x = p.x, xp = p.px * pn, y = p.y, yp = p.py * pn;
  const double g = aK * pn;
  const double t = sqrt(g);
  const double it = 1.0 / t;
  const double lt = L * t;
  double sn, cs;
  sincos(lt, &sn, &cs);
  const double ch = cosh(lt), sh = sinh(lt);
  double mhd, m12, m21, mvd, m34, m43;
  if (K > 0) {
    mhd = cs; m12 = sn * it; m21 = -m12 * g;
    mvd = ch; m34 = sh * it; m43 = m34 * g;
  } else {
    mhd = ch; m12 = sh * it; m21 = m12 * g;
    mvd = cs; m34 = sn * it; m43 = -m34 * g;
  }
  const double opd = 1.0 + p.d;
  p.x = mhd * x + m12 * xp;
  p.px = (m21 * x + mhd * xp) * opd;
  p.y = mvd * y + m34 * yp;
  p.py = (m43 * y + mvd * yp) * opd;
  p.z -= g * (x * x * (L - mhd * m12) - y * y * (L - mvd * m34)) * 0.25;
  p.z -= (xp * xp * (L + mhd * m12) + yp * yp * (L + mvd * m34)) * 0.25;
  p.z -= (x * xp * m12 * m21 + y * yp * m34 * m43) * 0.5;
}

// Hard-edge fringe with psi = irho*tan(edge) precomputed (atpass.cpp:190-196).
ACCP_HD void edge(Part &p, double psi) {
  p.px += p.x * psi;
  p.py -= p.y * psi;
}

// bend6 with ByError = 0 (atpass.cpp:76-187): linear sector bend with
// gradient; kx = angle/L.
ACCP_HD void bend_linear(Part &p, double L, double kx, double grd) {
  const double pn = p.ipd;
  const double g1 = (kx * kx + grd) * pn;
  const double g2 = -grd * pn;
  double mhd, m12, m21, mvd, m34, m43, s1 = 0, c1 = 1, sg1 = 0;
  if (g1 == 0) {
    mhd = 1; m12 = L; m21 = 0;
  } else if (g1 > 0) {
    sg1 = sqrt(g1);
    sincos(L * sg1, &s1, &c1);
    mhd = c1; m12 = s1 / sg1; m21 = -s1 * sg1;
  } else {
    sg1 = sqrt(-g1);
    const double a1 = L * sg1;
    s1 = sinh(a1); c1 = cosh(a1);
    mhd = c1; m12 = s1 / sg1; m21 = s1 * sg1;
  }
  if (g2 == 0) {
    mvd = 1; m34 = L; m43 = 0;
  } else if (g2 > 0) {
    const double sg2 = sqrt(g2);
    double s2, c2;
    sincos(L * sg2, &s2, &c2);
    mvd = c2; m34 = s2 / sg2; m43 = -s2 * sg2;
  } else {
    const double sg2 = sqrt(-g2);
    const double a2 = L * sg2;
    const double s2 = sinh(a2);
    mvd = cosh(a2); m34 = s2 / sg2; m43 = s2 * sg2;
  }
  const double x = p.x, xp = p.px * pn, y = p.y, yp = p.py * pn;
  const double opd = 1.0 + p.d;
  const double dp = p.d * pn;  // (delta*p_norm - ByError), ByError = 0
  double nx = mhd * x + m12 * xp;
  double npx = (m21 * x + mhd * xp) * opd;
  if (g1 == 0) {
    nx += dp * L * L * kx * 0.5;
    npx += dp * L * kx * opd;
  } else {
    nx += dp * (1 - c1) * kx / g1;
    npx += dp * s1 * kx / (sg1 * pn);
  }
  p.x = nx;
  p.px = npx;
  p.y = mvd * y + m34 * yp;
  p.py = (m43 * y + mvd * yp) * opd;
  double z = p.z - xp * xp * (L + mhd * m12) * 0.25;
  if (g1 != 0) {
    z -= (L - mhd * m12) * (x * x * g1 + dp * dp * kx * kx / g1 - 2 * x * kx * dp) * 0.25;
    z -= m12 * m21 * (x * xp - xp * dp * kx / g1) * 0.5;
    z -= kx * x * m12 + xp * (1 - mhd) * kx / g1 + dp * (L - m12) * kx * kx / g1;
  }
  z -= ((L - mvd * m34) * y * y * g2 + yp * yp * (L + mvd * m34)) * 0.25;
  z -= m34 * m43 * y * yp * 0.5;
  p.z = z;
}

// ATbendhxdrift6 (atpass.cpp:343-355); hsl4 = (1/4)*h*L*L precomputed,
// hs = h*L.
ACCP_HD void hxdrift(Part &p, double L, double h, double hs, double hsl4) {
  const double i1 = p.ipd;
  const double x = p.x, px = p.px, py = p.py;
  const double a = 1.0 + h * x;
  const double px2 = px * px, py2 = py * py;
  const double s = px2 + py2;
  p.x += a * px * i1 * L + hsl4 * (px2 - py2) * i1 * i1;
  p.px -= hs * s * i1 * 0.5;
  p.y += a * i1 * py * L * (1.0 + px * hs * 0.5);
  p.z -= a * i1 * i1 * L * 0.5 * s;
}

// Complex Horner of sum_{n<=mo} B[n] (x+iy)^n (atpass.cpp:224-229).
template <class PT>
ACCP_HD void horner_b(const PT *B, int mo, double x, double y, double &re, double &im) {
  double a = B[mo], b = 0.0;
  for (int i = mo - 1; i >= 0; --i) {
    const double t = a * x - b * y + B[i];
    b = b * x + a * y;
    a = t;
  }
  re = a;
  im = b;
}

// bndthinkick (atpass.cpp:199-235); k1c = B[1], k2c = B[2] if mo >= 2.
// `plain` = both curvature coefficients are zero (the terms vanish).
template <class PT>
ACCP_HD void bend_kick(Part &p, const PT *B, int mo, bool plain, double L, double h,
                       double k1c, double k2c) {
  const double x = p.x, y = p.y;
  double re, im;
  horner_b(B, mo, x, y, re, im);
  if (plain) {
    p.px -= L * (-h * p.d + re + h * (h * x));
    if (mo > 0) p.py += L * im;
  } else {
    p.px -= L * (-h * p.d + re + h * (h * x + k1c * (x * x - 0.5 * y * y) +
                                      k2c * (x * x * x - (4.0 / 3.0) * x * y * y)));
    p.py += L * (im + h * (k1c * x * y + (4.0 / 3.0) * k2c * x * x * y +
                           (h / 6.0 * k1c - k2c / 3.0) * y * y * y));
  }
  p.z -= L * h * x;
}

// strthinkick (atpass.cpp:358-414), Horner started at the highest non-zero
// order `hi` (leading zero coefficients contribute exact zeros).
template <class PT>
ACCP_HD void mpole_kick(Part &p, const PT *B, const PT *A, int hi, double L) {
  const double x = p.x, y = p.y;
  double re = B[hi], im = A[hi];
  for (int i = hi - 1; i >= 0; --i) {
    const double t = re * x - im * y + B[i];
    im = im * x + re * y + A[i];
    re = t;
  }
  p.px -= L * re;
  p.py += L * im;
}

// ACavity::pass (accel.cpp:903-923); w_c0 = 2*pi*f/C0.
ACCP_HD void cavity(Part &p, double L, double nv, double w_c0, double phis, bool on) {
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  if (on) set_delta(p, p.d + nv * sin(w_c0 * p.z + phis));
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// AAperture::pass (accel.cpp:337-387): loss code into z, x = NaN.
ACCP_HD bool ap_out(const Part &p, double x0, double x1, double y0, double y1) {
  return p.x <= x0 || p.x >= x1 || p.y <= y0 || p.y >= y1;
}
ACCP_HD void ap_mark(Part &p, double x0, double x1, double y0, double y1) {
  p.z = (double)((p.x <= x0) * 1 + (p.x >= x1) * 2 + (p.y <= y0) * 3 + (p.y >= y1) * 4);
  p.x = __builtin_nan("");
}
ACCP_HD void aperture(Part &p, double L, double x0, double x1, double y0, double y1) {
  if (ap_out(p, x0, x1, y0, y1)) {
    ap_mark(p, x0, x1, y0, y1);
    return;
  }
  if (L != 0) {
    drift_nl(p, L * p.ipd);
    if (ap_out(p, x0, x1, y0, y1)) ap_mark(p, x0, x1, y0, y1);
  }
}

// ACorrector::pass (accel.cpp:270-292).
ACCP_HD void corrector(Part &p, double L, double kx, double ky) {
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  p.px += kx * p.ipd;
  p.py += ky * p.ipd;
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// ---- fast paths: constants that depend only on delta hoisted per element ----
// Inside an element without radiation delta does not change, so 1/(1+delta)
// and everything built from it is computed once per particle per element
// instead of once per sub-step.

// fastdrift with z coefficient zc = NormL/(2(1+delta)).
ACCP_HD void drift_c(Part &p, double nl, double zc) {
  const double s = p.px * p.px + p.py * p.py;
  p.x += nl * p.px;
  p.y += nl * p.py;
  p.z -= zc * s;
}

// Per-particle constants of ATbendhxdrift6 for one step length L:
// c1 = L/(1+d), c2 = (h L^2/4)/(1+d)^2, c3 = (h L/2)/(1+d), c5 = L/(2(1+d)^2);
// c4 = h L/2 is particle independent.
struct HxC {
  double c1, c2, c3, c5;
};
ACCP_HD HxC hx_consts(double ipd, double L, double hs, double hsl4) {
  return HxC{ipd * L, hsl4 * ipd * ipd, hs * 0.5 * ipd, ipd * ipd * L * 0.5};
}
ACCP_HD void hxdrift_c(Part &p, const HxC &c, double h, double c4) {
  const double x = p.x, px = p.px, py = p.py;
  const double a = 1.0 + h * x;
  const double px2 = px * px, py2 = py * py;
  const double s = px2 + py2;
  p.x += a * px * c.c1 + c.c2 * (px2 - py2);
  p.px -= c.c3 * s;
  p.y += a * py * c.c1 * (1.0 + px * c4);
  p.z -= a * c.c5 * s;
}

// ATbendhxdrift6 in a bend whose kicks leave py alone (no field above the
// dipole term, no curvature gradients): py is fixed through the slices, so
// its terms are per-particle constants of the element.  With
// g = c1 (1 + h x) (the reference's a*i1*L) the update regroups as
//   x  += px (g + c2 px) - c2 py^2
//   px -= c3 (px^2 + py^2)
//   y  += g (py + c4 py px)
//   z  -= (1/(2(1+delta))) g (px^2 + py^2)
// with c1..c4 as in HxC.  Neither the pseudo-drifts nor the kicks read z,
// so its increments are summed in w (in units of 2(1+delta), where each
// pseudo-drift adds g s in one fused operation) and added to z once at the
// end of the element (dipole_close).  The reference's terms in another
// order: round-off differences only, 9 operations per pseudo-drift where
// the reference's order takes 11.
struct HxW {
  double c1, c1h, c2, c3, c2py2, pyc4;
};
ACCP_HD HxW hx_consts_w(double ipd, double L, double hs, double hsl4, double py, double h) {
  const double c1 = ipd * L, c2 = hsl4 * ipd * ipd;
  return HxW{c1, c1 * h, c2, hs * 0.5 * ipd, c2 * (py * py), py * (hs * 0.5)};
}
ACCP_HD void hxdrift_w(Part &p, const HxW &c, double py, double py2, double &w) {
  const double x = p.x, px = p.px;
  const double g = fma(c.c1h, x, c.c1);
  const double s = fma(px, px, py2);
  p.x = fma(px, fma(c.c2, px, g), x - c.c2py2);
  p.px = fma(-c.c3, s, px);
  p.y = fma(g, fma(c.pyc4, px, py), p.y);
  w = fma(g, s, w);
}
// bndthinkick with PolyB[0] alone and delta fixed: px -= L (B0 + h h x) -
// L h delta, as px + q - (L h h) x with q = L h delta - L B0 per particle
// and element; z -= L h x, as w += lw x with lw = 2 (1+delta) L h
// (atpass.cpp:229-233).
ACCP_HD void bend_kick_w(Part &p, double lhh, double q, double lw, double &w) {
  const double x = p.x;
  p.px = fma(-lhh, x, p.px + q);
  w = fma(lw, x, w);
}
// z after the element: z - w / (2 (1+delta)).
ACCP_HD void dipole_close(Part &p, double w) { p.z = fma(w, -0.5 * p.ipd, p.z); }
// The kick of a pure sextupole term b (x + i y)^2: px -= k b (x^2 - y^2),
// py += 2 k b x y, with kb = k b and kb2 = 2 k b (the reference's Horner sum
// in another order: round-off differences only).
ACCP_HD void sext_kick(Part &p, double kb, double kb2) {
  const double x = p.x, y = p.y;
  p.px = fma(-kb, fma(x, x, -(y * y)), p.px);
  p.py = fma(kb2 * x, y, p.py);
}

// The ns Forest-Ruth slices of a pure sextupole (AStrMPole::pass with
// PolyB[2] alone, accel.cpp:829-843): drift(n1) kick(k1) drift(n2) kick(k2)
// drift(n2) kick(k1) drift(n1) each, the kicks by sext_kick.  The last drift
// of one slice and the first of the next are one drift of twice the length:
// the same map in exact arithmetic (px, py do not change in a drift), rounded
// once instead of twice, and 9 drifts fewer in 10 slices.
ACCP_HD void sext_slice_inner(Part &p, double n2, double z2, double kb1, double kb1x2, double kb2, double kb2x2) {
  sext_kick(p, kb1, kb1x2);
  drift_c(p, n2, z2);
  sext_kick(p, kb2, kb2x2);
  drift_c(p, n2, z2);
  sext_kick(p, kb1, kb1x2);
}
ACCP_HD void sext_slices(Part &p, int ns, double n1, double z1, double n2, double z2, double kb1, double kb1x2,
                         double kb2, double kb2x2) {
  if (ns <= 0) return;
  const double n1d = 2.0 * n1, z1d = 2.0 * z1;
  drift_c(p, n1, z1);
  if (ns == 10) {  // the reference's default nslice (accel.h:242): straight-line code
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      sext_slice_inner(p, n2, z2, kb1, kb1x2, kb2, kb2x2);
      drift_c(p, n1d, z1d);
    }
  } else {
    for (int m = 0; m < ns - 1; ++m) {
      sext_slice_inner(p, n2, z2, kb1, kb1x2, kb2, kb2x2);
      drift_c(p, n1d, z1d);
    }
  }
  sext_slice_inner(p, n2, z2, kb1, kb1x2, kb2, kb2x2);
  drift_c(p, n1, z1);
}

// Complex Horner of order HI over coefficients held in registers.
template <int HI>
ACCP_HD void horner_reg(const double (&b)[HI + 1], const double (&a)[HI + 1], double x, double y,
                        double &re, double &im) {
  re = b[HI];
  im = a[HI];
#pragma unroll
  for (int i = HI - 1; i >= 0; --i) {
    const double t = re * x - im * y + b[i];
    im = im * x + re * y + a[i];
    re = t;
  }
}
// The same sum of order 9, sum_n (b_n + i a_n)(x + i y)^n, by Estrin's scheme:
// pairs c_2k + c_2k+1 z, then z^2, z^4, z^8 combine them in three levels, a
// dependent chain of 8 fused operations where Horner's has 18 (the terms in
// another order: round-off differences only).
ACCP_HD void estrin9(const double (&b)[10], const double (&a)[10], double x, double y, double &re, double &im) {
  double er[5], ei[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    er[k] = fma(b[2 * k + 1], x, fma(-a[2 * k + 1], y, b[2 * k]));
    ei[k] = fma(b[2 * k + 1], y, fma(a[2 * k + 1], x, a[2 * k]));
  }
  const double z2r = fma(x, x, -(y * y)), z2i = (x + x) * y;
  const double z4r = fma(z2r, z2r, -(z2i * z2i)), z4i = (z2r + z2r) * z2i;
  const double z8r = fma(z4r, z4r, -(z4i * z4i)), z8i = (z4r + z4r) * z4i;
  const double f0r = fma(er[1], z2r, fma(-ei[1], z2i, er[0])), f0i = fma(er[1], z2i, fma(ei[1], z2r, ei[0]));
  const double f1r = fma(er[3], z2r, fma(-ei[3], z2i, er[2])), f1i = fma(er[3], z2i, fma(ei[3], z2r, ei[2]));
  const double g0r = fma(f1r, z4r, fma(-f1i, z4i, f0r)), g0i = fma(f1r, z4i, fma(f1i, z4r, f0i));
  re = fma(er[4], z8r, fma(-ei[4], z8i, g0r));
  im = fma(er[4], z8i, fma(ei[4], z8r, g0i));
}
// The order-9 sum with c_0 = c_1 = 0 (c_n = b_n + i a_n):
// z^2 sum_{m<=7} c_{m+2} z^m, the inner sum by Estrin's scheme (pairs,
// then z^2 and z^4), then one complex product by z^2: 40 operations where
// estrin9 takes 48 (the terms in another order: round-off differences only).
ACCP_HD void estrin9_lo2(const double (&b)[10], const double (&a)[10], double x, double y, double &re,
                         double &im) {
  double er[4], ei[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    er[k] = fma(b[2 * k + 3], x, fma(-a[2 * k + 3], y, b[2 * k + 2]));
    ei[k] = fma(b[2 * k + 3], y, fma(a[2 * k + 3], x, a[2 * k + 2]));
  }
  const double z2r = fma(x, x, -(y * y)), z2i = (x + x) * y;
  const double z4r = fma(z2r, z2r, -(z2i * z2i)), z4i = (z2r + z2r) * z2i;
  const double f0r = fma(er[1], z2r, fma(-ei[1], z2i, er[0])), f0i = fma(er[1], z2i, fma(ei[1], z2r, ei[0]));
  const double f1r = fma(er[3], z2r, fma(-ei[3], z2i, er[2])), f1i = fma(er[3], z2i, fma(ei[3], z2r, ei[2]));
  const double gr = fma(f1r, z4r, fma(-f1i, z4i, f0r)), gi = fma(f1r, z4i, fma(f1i, z4r, f0i));
  re = fma(gr, z2r, -(gi * z2i));
  im = fma(gr, z2i, gi * z2r);
}
// b (x+iy)^HI: horner_reg on the coefficients {0, ..., 0, b} and A = 0 with
// the additions of those zeros and the first step's products by the zero
// imaginary part left out.  For finite coordinates each is an exact zero, so
// the sum is the full one's; only an orbit that has already overflowed to
// infinity can see a finite value here where the full sum has 0 * inf = NaN
// (its x is infinite or NaN by then, so its loss is the same).
template <int HI>
ACCP_HD void horner_single(double b, double x, double y, double &re, double &im) {
  re = b * x;
  im = b * y;
#pragma unroll
  for (int i = HI - 2; i >= 0; --i) {
    const double t = re * x - im * y;
    im = im * x + re * y;
    re = t;
  }
}
template <int HI>
ACCP_HD void horner_b_reg(const double (&b)[HI + 1], double x, double y, double &re, double &im) {
  re = b[HI];
  im = 0.0;
#pragma unroll
  for (int i = HI - 1; i >= 0; --i) {
    const double t = re * x - im * y + b[i];
    im = im * x + re * y;
    re = t;
  }
}

// bndthinkick with delta fixed: lhd = L*h*delta precomputed per element.
template <int MO, bool PLAIN>
ACCP_HD void bend_kick_c(Part &p, const double (&b)[MO + 1], double L, double h, double lhd,
                         double lh, double k1c, double k2c) {
  const double x = p.x, y = p.y;
  double re, im;
  horner_b_reg<MO>(b, x, y, re, im);
  if (PLAIN) {
    p.px -= L * (re + h * (h * x)) - lhd;
    if (MO > 0) p.py += L * im;
  } else {
    p.px -= L * (re + h * (h * x + k1c * (x * x - 0.5 * y * y) +
                           k2c * (x * x * x - (4.0 / 3.0) * x * y * y))) - lhd;
    p.py += L * (im + h * (k1c * x * y + (4.0 / 3.0) * k2c * x * x * y +
                           (h / 6.0 * k1c - k2c / 3.0) * y * y * y));
  }
  p.z -= lh * x;
}

// quad6 for K != 0 from power series in u = L^2 |K|/(1+delta):
//   cos(sqrt u) = E(-u), cosh(sqrt u) = E(u),       E(v) = sum v^k/(2k)!
//   sin(sqrt u)/sqrt u = O(-u), sinh(sqrt u)/sqrt u = O(u),  O(v) = sum v^k/(2k+1)!
// split into even/odd k so both signs come from two polynomials in u^2:
// no sqrt, division or transcendental.  NT terms are exact to double
// rounding for u <= umax chosen by the host.
// C holds the series coefficients in four blocks of M = NT/2 values:
// 1/(4m)!, 1/(4m+2)!, 1/(4m+1)!, 1/(4m+3)! for m = 0..M-1 (host-computed, so
// no constant lives in registers across the kernel).
template <int NT, class PT>
ACCP_HD void cs_series(const PT *C, double u, double &c, double &ch, double &s, double &sh) {
  constexpr int M = NT / 2;
  const double w = u * u;
  double ee = C[M - 1], eo = C[2 * M - 1], oe = C[3 * M - 1], oo = C[4 * M - 1];
#pragma unroll
  for (int m = M - 2; m >= 0; --m) {
    ee = ee * w + C[m];
    eo = eo * w + C[M + m];
    oe = oe * w + C[2 * M + m];
    oo = oo * w + C[3 * M + m];
  }
  c = ee - u * eo;
  ch = ee + u * eo;
  s = oe - u * oo;
  sh = oe + u * oo;
}

template <int NT, class PT>
ACCP_HD void quad_series(Part &p, const PT *C, double L, double L2, bool focusing, double aK) {
  const double pn = p.ipd;
  const double x = p.x, xp = p.px * pn, y = p.y, yp = p.py * pn;
  const double g = aK * pn;
  double c, ch, s, sh;
  cs_series<NT>(C, L2 * g, c, ch, s, sh);
  double mhd, m12, m21, mvd, m34, m43;
  if (focusing) {
    mhd = c; m12 = L * s; m21 = -m12 * g;
    mvd = ch; m34 = L * sh; m43 = m34 * g;
  } else {
    mhd = ch; m12 = L * sh; m21 = m12 * g;
    mvd = c; m34 = L * s; m43 = -m34 * g;
  }
  const double opd = 1.0 + p.d;
  p.x = mhd * x + m12 * xp;
  p.px = (m21 * x + mhd * xp) * opd;
  p.y = mvd * y + m34 * yp;
  p.py = (m43 * y + mvd * yp) * opd;
  p.z -= g * (x * x * (L - mhd * m12) - y * y * (L - mvd * m34)) * 0.25;
  p.z -= (xp * xp * (L + mhd * m12) + yp * yp * (L + mvd * m34)) * 0.25;
  p.z -= (x * xp * m12 * m21 + y * yp * m34 * m43) * 0.5;
}

// ---- AMatrix: 6x6 map between half drifts (accel.cpp:1697-1729) -------------
template <class PT>
ACCP_HD void matrix_map(Part &p, double L, const PT *m) {
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  const double r[6] = {p.x, p.px, p.y, p.py, p.z, p.d};
  double t[6];
#pragma unroll
  for (int j = 0; j < 6; ++j)
    t[j] = m[j * 6 + 0] * r[0] + m[j * 6 + 1] * r[1] + m[j * 6 + 2] * r[2] + m[j * 6 + 3] * r[3] +
           m[j * 6 + 4] * r[4] + m[j * 6 + 5] * r[5];
  p.x = t[0];
  p.px = t[1];
  p.y = t[2];
  p.py = t[3];
  p.z = t[4];
  set_delta(p, t[5]);
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// ---- bndthinkickrad (atpass.cpp:255-339) -------------------------------------
// crad = CGAMMA*E0^3/(2*pi*1e9); cq = 55*sqrt(3)/72*uc/E0; flag 1 = classical
// loss, flag 2 = loss plus gauss*sigma (quantum excitation).  The curvature
// (Iselin) terms are added to the multipole sum before it enters both the
// radiated power and the kick.  rcp(v) returns 1/v for the velocity norm
// (atpass.cpp:242) and for 1/(1+delta) after the loss.  PLAIN: k1c = k2c =
// 0, and the curvature terms, exact zeros then, are left out (the same
// values up to the sign of a zero).
template <bool PLAIN = false, class PT, class Rcp>
ACCP_HD void bend_kick_rad(Part &p, const PT *B, int mo, double L, double h, double k1c, double k2c,
                           double crad, double cq, int flag, double gauss, Rcp rcp) {
  const double pn = p.ipd;
  const double x = p.x, xp = p.px * pn, y = p.y, yp = p.py * pn;
  double re, im;
  horner_b(B, mo, x, y, re, im);
  if constexpr (!PLAIN) {
    const double hk = h * (k1c * h - k2c);
    im += hk * y * y * y * (1.0 / 6.0);
    re += -k1c * h * y * y * 0.5 + hk * x * y * y * 0.5;
  }
  const double a = 1.0 + x * h;
  const double vn2 = rcp(a * a + xp * xp + yp * yp);
  const double by = re + h, bx = im;
  const double t1 = by * a, t2 = bx * a, t3 = bx * yp - by * xp;
  const double b2p = (t1 * t1 + t2 * t2 + t3 * t3) * vn2;
  const double opd = 1.0 + p.d;
  const double mean = crad * (opd * opd) * b2p * (a + (xp * xp + yp * yp) * 0.5) * L;
  double dnew;
  if (flag == 1) {
    dnew = p.d - mean;
  } else {
    const double sig = sqrt(cq * fabs(mean));
    dnew = p.d - (mean + gauss * sig);
  }
  const double opn = 1.0 + dnew;
  p.d = dnew;
  p.ipd = rcp(opn);
  p.px = xp * opn;
  p.py = yp * opn;
  if constexpr (PLAIN) {
    p.px -= L * (-h * dnew + re + h * (h * x));
    p.py += L * im;
  } else {
    p.px -= L * (-h * dnew + re + h * (h * x + k1c * (x * x - 0.5 * y * y) +
                                       k2c * (x * x * x - (4.0 / 3.0) * x * y * y)));
    p.py += L * (im + h * (k1c * x * y + (4.0 / 3.0) * k2c * x * x * y +
                           (h / 6.0 * k1c - k2c / 3.0) * y * y * y));
  }
  p.z -= L * h * x;
}

// ---- ASyncRad: lumped damping and excitation (accel.cpp:1813-1854) ----------
// C = [L, 1-ax, 1-ay, 1-az, sigx sx, sigy sy, sigdpp sz]; g = three normals.
template <class PT>
ACCP_HD void syncrad(Part &p, const PT *C, bool damp, bool qe, double g0, double g1, double g2) {
  const double L = C[0];
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  double d = p.d;
  if (damp) {
    p.x *= C[1];
    p.y *= C[2];
    d *= C[3];
  }
  if (qe) {
    p.x += C[4] * g0;
    p.y += C[5] * g1;
    d += C[6] * g2;
  }
  set_delta(p, d);
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// ---- ACrabCav: transverse deflecting cavity (accel.cpp:968-999) -------------
// C = [L, Vx/E, Vy/E, k, phis, sigphi, sigdVV]; g = three normals.
template <class PT>
ACCP_HD void crabcav(Part &p, const PT *C, int flags, double g0, double g1, double g2) {
  const double L = C[0];
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  if (flags & 1) {
    const double k = C[3];
    double om = -k * p.z + C[4];
    if (flags & 2) om = -k * p.z + C[4] + C[5] * g0;
    double nvxa = C[1], nvya = C[2];
    if (flags & 4) {
      nvxa = C[1] * (1.0 + C[6] * g1);
      nvya = C[2] * (1.0 + C[6] * g2);
    }
    const double so = sin(om), co = cos(om);
    p.px += nvxa * so;
    p.py += nvya * so;
    set_delta(p, p.d + -(nvya * p.y + nvxa * p.x) * k * co / (1 + p.d));
  }
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// ---- ARFMult: RF multipole (accel.cpp:1146-1201) ---------------------------
// C = [L, k, phis, sigphi, B0..Bn-1, A0..An-1]; g = phase normal.
template <class PT>
ACCP_HD void rfmult(Part &p, const PT *C, int n, bool noise, double g) {
  const double L = C[0], k = C[1];
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
  const double om = noise ? -k * p.z + C[2] + C[3] * g : -k * p.z + C[2];
  const double rho = sqrt(p.x * p.x + p.y * p.y);
  const double ph = atan2(p.y, p.x);
  const PT *B = C + 4, *A = C + 4 + n;
  double AsL = 0, ArhoL = 0, AphiL = 0, nrho = 1.0;
  for (int m = 0; m < n; ++m) {
    double sn, cs;
    sincos((m + 1) * ph, &sn, &cs);
    const double td = B[m] * cs + A[m] * sn;
    ArhoL += nrho * td;
    AphiL += nrho * (-B[m] * sn + A[m] * cs);
    nrho *= rho;
    AsL += 1.0 / (m + 1) * nrho * td;
  }
  double sph, cph, so, co;
  sincos(ph, &sph, &cph);
  sincos(om, &so, &co);
  p.px += (ArhoL * cph - AphiL * sph) * co;
  p.py += (ArhoL * sph + AphiL * cph) * co;
  set_delta(p, p.d + k * AsL * so / (1. + p.d));
  if (L != 0) drift_nl(p, 0.5 * L * p.ipd);
}

// ---- AIdealWig: ideal wiggler (accel.cpp:1463-1537) -------------------------
// C = [L, k^2, h^2, L/nslice, L/(2 nslice), crad h^2 L/nslice/2].
template <class PT>
ACCP_HD void idealwig(Part &p, const PT *C, int nslice, int type, bool rad) {
  const double k2 = C[1], hh = C[2], lns = C[3], half = C[4], rq = C[5];
  for (int j = 0; j < nslice; ++j) {
    drift_nl(p, half * p.ipd);
    const double y = p.y, x = p.x;
    if (rad) {  // half of the slice's classical loss, angles kept
      const double xp = p.px * p.ipd, yp = p.py * p.ipd;
      set_delta(p, p.d - rq * (1 + p.d) * (1 + p.d));
      p.px = xp * (1 + p.d);
      p.py = yp * (1 + p.d);
    }
    if (type == 0) {
      const double y3 = y * y * y;
      p.py -= hh * (0.5 * y + 1.0 / 3.0 * y3 * k2 + 1.0 / 15.0 * y3 * y * y * k2 * k2) * lns;
    } else if (type == 1) {
      const double x3 = x * x * x;
      p.px -= hh * (0.5 * x + 1.0 / 3.0 * x3 * k2 + 1.0 / 15.0 * x3 * x * x * k2 * k2) * lns;
    }
    if (rad) {
      const double xp = p.px * p.ipd, yp = p.py * p.ipd;
      set_delta(p, p.d - rq * (1 + p.d) * (1 + p.d));
      p.px = xp * (1 + p.d);
      p.py = yp * (1 + p.d);
    }
    drift_nl(p, half * p.ipd);
  }
}

// ---- AWigTable: kick table with bilinear interpolation (accel.cpp:1338-1382)
// Interval of a monotonic grid by bisection, the lower index clamped to
// [0, n-2] so that points outside the grid extrapolate from the edge cell
// (interp2, utility.cpp:174-187, through a fresh bilinear interpolator per
// call, whose first lookup is always the bisection).
template <class PT>
ACCP_HD int wt_cell(const PT *g, int n, double v) {
  const bool asc = g[n - 1] >= g[0];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (hi + lo) >> 1;
    if ((v >= g[mid]) == asc) lo = mid;
    else hi = mid;
  }
  return lo < n - 2 ? lo : n - 2;
}

// Bilinear value of tab (my rows of mx) at (x, y); t and u are the cell
// fractions, the four corners summed in the interpolator's order.
template <class PT>
ACCP_HD double wt_interp(const PT *xl, int mx, const PT *yl, int my, const PT *tab, double x, double y) {
  const int i = wt_cell(xl, mx, x), j = wt_cell(yl, my, y);
  const double t = (x - xl[i]) / (xl[i + 1] - xl[i]);
  const double u = (y - yl[j]) / (yl[j + 1] - yl[j]);
  const PT *r0 = tab + j * mx + i, *r1 = r0 + mx;
  return (1. - t) * (1. - u) * r0[0] + t * (1. - u) * r0[1] + (1. - t) * u * r1[0] + t * u * r1[1];
}

// C = [L/2, nslice, xl[mx], yl[my], xkick[my][mx], ykick[my][mx]];
// nslice = 0: one full kick (zero length).
template <class PT>
ACCP_HD void wigtable(Part &p, const PT *C, int nslice, int mx, int my) {
  const PT *xl = C + 2, *yl = xl + mx, *xk = yl + my, *yk = xk + mx * my;
  if (nslice == 0) {
    const double x = p.x, y = p.y;
    p.px += wt_interp(xl, mx, yl, my, xk, x, y);
    p.py += wt_interp(xl, mx, yl, my, yk, x, y);
    return;
  }
  const double ns = C[1];
  const double nl = C[0] * p.ipd / ns;  // 0.5 L/(1+delta)/nslice
  for (int s = 0; s < nslice; ++s) {
    drift_nl(p, nl);
    const double x = p.x, y = p.y;
    p.px += wt_interp(xl, mx, yl, my, xk, x, y) / ns;
    p.py += wt_interp(xl, mx, yl, my, yk, x, y) / ns;
    drift_nl(p, nl);
  }
}

// ---- counter-based normal variates (quantum excitation) -------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with a
// Weyl-sequence key schedule.
ACCP_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

}  // namespace accp
