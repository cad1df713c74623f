/*
 * accp_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the Accel+ tracking arithmetic.  Each function keeps
 * the reference's operation order so that, compiled with gcc -O2
 * -ffp-contract=off and no -march (no FMA on x86-64), it reproduces the
 * reference bit for bit; the tests check that against fixtures written by
 * the compiled reference.  Citations are file:line in the Accel+ tree.
 */
#include "accp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.1415926535897932384626       /* accel.h:23 */
#define ORC_TWOPI (2 * ORC_PI)                 /* accel.h:24 */
#define ORC_C0 2.99792458e8                    /* accel.h:29 */
#define ORC_CGAMMA 8.846056192e-05             /* atpass.cpp:270 */
/* Forest-Ruth coefficients, accel.h:44-47 (decimal literals kept). */
#define ORC_DRIFT1 0.6756035959798286638
#define ORC_DRIFT2 -0.1756035959798286639
#define ORC_KICK1 1.351207191959657328
#define ORC_KICK2 -1.702414383919314656
#define ORC_DELTA_DIFF 1.0e-8                  /* accel.h:586 */

static inline double sq(double v) { return v * v; }

/* atpass.cpp:6-13 -- only reached from quad6 with K == 0.  The reference
 * squares (px + y^2) instead of px^2 + py^2; that is reproduced here. */
void orc_drift_kzero(double *r, double L) {
  double pn = 1.0 / (1 + r[5]);
  r[0] += r[1] * L;
  r[2] += r[3] * L;
  r[4] -= sq(pn) * L * (sq(r[1] + sq(r[2]))) / 2.0;
}

/* atpass.cpp:15-72 -- linear chromatic quadrupole with second-order path
 * length.  g = |K|/(1+delta); focusing plane cos/sin, other plane cosh/sinh. */
void orc_quad6(double *r, double L, double K) {
  if (K == 0) {
    orc_drift_kzero(r, L);
    return;
  }
  double pn = 1 / (1 + r[5]);
  double x = r[0], xp = r[1] * pn, y = r[2], yp = r[3] * pn;
  double g = fabs(K) / (1 + r[5]);
  double t = sqrt(g);
  double lt = L * t;
  double cx, cy, c12, c21, c34, c43;
  if (K > 0) {
    cx = cos(lt);
    c12 = sin(lt) / t;
    c21 = -c12 * g;
    cy = cosh(lt);
    c34 = sinh(lt) / t;
    c43 = c34 * g;
  } else {
    cx = cosh(lt);
    c12 = sinh(lt) / t;
    c21 = c12 * g;
    cy = cos(lt);
    c34 = sin(lt) / t;
    c43 = -c34 * g;
  }
  r[0] = cx * x + c12 * xp;
  r[1] = (c21 * x + cx * xp) / pn;
  r[2] = cy * y + c34 * yp;
  r[3] = (c43 * y + cy * yp) / pn;
  r[4] -= g * (x * x * (L - cx * c12) - y * y * (L - cy * c34)) / 4;
  r[4] -= (xp * xp * (L + cx * c12) + yp * yp * (L + cy * c34)) / 4;
  r[4] -= (x * xp * c12 * c21 + y * yp * c34 * c43) / 2;
}

/* atpass.cpp:76-187 -- linear sector bend with gradient. */
void orc_bend6(double *r, double L, double angle, double grd, double by_err) {
  double m12, m21, m34, m43, mvd, mhd;
  double sg1 = 0, sg2, a1 = 0, a2;
  double kx = angle / L;
  double pn = 1 / (1 + r[5]);
  double g1 = (kx * kx + grd) * pn;
  double g2 = -grd * pn;
  if (g1 == 0) {
    mhd = 1; m12 = L; m21 = 0;
  } else if (g1 > 0) {
    sg1 = sqrt(g1); a1 = L * sg1;
    mhd = cos(a1); m12 = sin(a1) / sg1; m21 = -sin(a1) * sg1;
  } else {
    sg1 = sqrt(-g1); a1 = L * sg1;
    mhd = cosh(a1); m12 = sinh(a1) / sg1; m21 = sinh(a1) * sg1;
  }
  if (g2 == 0) {
    mvd = 1; m34 = L; m43 = 0;
  } else if (g2 > 0) {
    sg2 = sqrt(g2); a2 = L * sg2;
    mvd = cos(a2); m34 = sin(a2) / sg2; m43 = -sin(a2) * sg2;
  } else {
    sg2 = sqrt(-g2); a2 = L * sg2;
    mvd = cosh(a2); m34 = sinh(a2) / sg2; m43 = sinh(a2) * sg2;
  }
  double x = r[0], xp = r[1] * pn, y = r[2], yp = r[3] * pn, d = r[5];
  r[0] = mhd * x + m12 * xp;
  r[1] = (m21 * x + mhd * xp) / pn;
  if (g1 == 0) {
    r[0] += (d * pn - by_err) * L * L * kx / 2;
    r[1] += (d * pn - by_err) * L * kx / pn;
  } else if (g1 > 0) {
    r[0] += (d * pn - by_err) * (1 - cos(a1)) * kx / g1;
    r[1] += (d * pn - by_err) * sin(a1) * kx / (sg1 * pn);
  } else {
    r[0] += (d * pn - by_err) * (1 - cosh(a1)) * kx / g1;
    r[1] += (d * pn - by_err) * sinh(a1) * kx / (sg1 * pn);
  }
  r[2] = mvd * y + m34 * yp;
  r[3] = (m43 * y + mvd * yp) / pn;
  r[4] -= xp * xp * (L + mhd * m12) / 4;
  if (g1 != 0) {
    r[4] -= (L - mhd * m12) * (x * x * g1 + (d * pn - by_err) * (d * pn - by_err) * kx * kx / g1
                               - 2 * x * kx * (d * pn - by_err)) / 4;
    r[4] -= m12 * m21 * (x * xp - xp * (d * pn - by_err) * kx / g1) / 2;
    r[4] -= kx * x * m12 + xp * (1 - mhd) * kx / g1 + (d * pn - by_err) * (L - m12) * kx * kx / g1;
  }
  r[4] -= ((L - mvd * m34) * y * y * g2 + yp * yp * (L + mvd * m34)) / 4;
  r[4] -= m34 * m43 * y * yp / 2;
}

/* atpass.cpp:190-196 -- hard-edge dipole fringe. */
void orc_edge(double *r, double inv_rho, double edge_angle) {
  double psi = inv_rho * tan(edge_angle);
  r[1] += r[0] * psi;
  r[3] -= r[2] * psi;
}

/* Complex Horner of sum B[n] (x+iy)^n (normal part only), from max_order
 * down; the loop shared by bndthinkick and bndthinkickrad
 * (atpass.cpp:224-229, :288-293). */
static void horner_b(const double *B, int max_order, double x, double y,
                     double *re, double *im) {
  double a = B[max_order], b = 0;
  for (int i = max_order - 1; i >= 0; i--) {
    double t = a * x - b * y + B[i];
    b = b * x + a * y;
    a = t;
  }
  *re = a;
  *im = b;
}

/* atpass.cpp:199-235 -- thin kick in the curved frame. */
void orc_bndthinkick(double *r, const double *B, double L, double irho, int max_order) {
  double k1 = B[1];
  double k2 = max_order >= 2 ? B[2] : 0;
  double h = irho, re, im;
  horner_b(B, max_order, r[0], r[2], &re, &im);
  r[1] -= L * (-h * r[5] + re + h * (h * r[0] + k1 * (r[0] * r[0] - 0.5 * r[2] * r[2])
                                      + k2 * (r[0] * r[0] * r[0] - 4.0 / 3.0 * r[0] * r[2] * r[2])));
  r[3] += L * (im + h * (k1 * r[0] * r[2] + 4.0 / 3.0 * k2 * r[0] * r[0] * r[2]
                          + (h / 6.0 * k1 - k2 / 3.0) * r[2] * r[2] * r[2]));
  r[4] -= L * h * r[0];
}

/* atpass.cpp:237-253 -- |e x B|^2 with e the unit velocity. */
double orc_b2perp(double bx, double by, double irho, double x, double xpr,
                  double y, double ypr) {
  (void)y;
  double vn2 = 1 / (sq(1 + x * irho) + sq(xpr) + sq(ypr));
  return (sq(by * (1 + x * irho)) + sq(bx * (1 + x * irho)) + sq(bx * ypr - by * xpr)) * vn2;
}

/* atpass.cpp:255-339 -- thin kick with classical radiation (flag 1) or
 * radiation plus quantum excitation (flag 2).  `gauss` is the standard
 * normal used where the reference calls grandn() (atpass.cpp:323). */
void orc_bndthinkickrad(double *r, const double *B, double L, double irho,
                        double E0, int max_order, int flag_rad, double gauss) {
  double crad = ORC_CGAMMA * E0 * E0 * E0 / (ORC_TWOPI * 1e9);
  double uc_e0 = 0.665 * E0 * E0 / 1.0e6 / 2.99792458E8 * fabs(irho) * 1000;
  double k1 = B[1];
  double k2 = max_order >= 2 ? B[2] : 0;
  double h = irho, re, im;
  horner_b(B, max_order, r[0], r[2], &re, &im);
  double pn = 1 / (1 + r[5]);
  double x = r[0], xp = r[1] * pn, y = r[2], yp = r[3] * pn;
  im += h * (k1 * h - k2) * y * y * y / 6.0;
  re += -k1 * h * y * y / 2.0 + h * (k1 * h - k2) * x * y * y / 2.0;
  double b2p = orc_b2perp(im, re + irho, irho, x, xp, y, yp);
  if (flag_rad == 1) {
    r[5] = r[5] - crad * sq(1 + r[5]) * b2p * (1 + x * irho + (sq(xp) + sq(yp)) / 2) * L;
  } else if (flag_rad == 2) {
    double mean = crad * sq(1 + r[5]) * b2p * (1 + x * irho + (sq(xp) + sq(yp)) / 2) * L;
    double sig;
    if (mean >= 0)
      sig = sqrt(55 * sqrt(3) / 72.0 * uc_e0 * mean);
    else
      sig = sqrt(-55 * sqrt(3) / 72.0 * uc_e0 * mean);
    double dd = mean + gauss * sig;
    r[5] = r[5] - dd;
  }
  pn = 1 / (1 + r[5]);
  r[1] = xp / pn;
  r[3] = yp / pn;
  r[1] -= L * (-h * r[5] + re + h * (h * r[0] + k1 * (r[0] * r[0] - 0.5 * r[2] * r[2])
                                      + k2 * (r[0] * r[0] * r[0] - 4.0 / 3.0 * r[0] * r[2] * r[2])));
  r[3] += L * (im + h * (k1 * r[0] * r[2] + 4.0 / 3.0 * k2 * r[0] * r[0] * r[2]
                          + (h / 6.0 * k1 - k2 / 3.0) * r[2] * r[2] * r[2]));
  r[4] -= L * h * r[0];
}

/* atpass.cpp:343-355 -- pseudo-drift of H1 = (1+hx)(px^2+py^2)/2(1+delta). */
void orc_hxdrift6(double *r, double L, double h) {
  double hs = h * L;
  double i1pd = 1.0 / (1 + r[5]);
  double x = r[0], px = r[1], y = r[2], py = r[3];
  (void)y;
  r[0] += (1 + h * x) * px * i1pd * L + 1 / 4. * hs * L * (px * px - py * py) * i1pd * i1pd;
  r[1] -= hs * (px * px + py * py) * i1pd / 2.0;
  r[2] += (1.0 + h * x) * i1pd * py * L * (1. + px * hs / 2.0);
  r[4] -= (1.0 + h * x) * i1pd * i1pd * L / 2.0 * (px * px + py * py);
}

/* atpass.cpp:358-414 -- straight multipole kick, Horner over (B + iA). */
void orc_strthinkick(double *r, const double *A, const double *B, double L, int max_order) {
  double re = B[max_order], im = A[max_order];
  for (int i = max_order - 1; i >= 0; i--) {
    double t = re * r[0] - im * r[2] + B[i];
    im = im * r[0] + re * r[2] + A[i];
    re = t;
  }
  r[1] -= L * re;
  r[3] += L * im;
}

/* atpass.cpp:416-428 -- drift with precomputed L/(1+delta). */
void orc_fastdrift(double *r, double norm_l) {
  double dx = norm_l * r[1];
  double dy = norm_l * r[3];
  r[0] += dx;
  r[2] += dy;
  r[4] -= norm_l * (r[1] * r[1] + r[3] * r[3]) / (2 * (1 + r[5]));
}

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in the open interval (0,1) from two 32-bit words. */
static double u53(uint32_t a, uint32_t b) {
  uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)k + 0.5) * (1.0 / 9007199254740992.0);
}

double orc_gauss(uint64_t seed, uint64_t pid, uint64_t turn, uint32_t eidx, uint32_t slice) {
  uint32_t ctr[4] = {(uint32_t)pid, (uint32_t)turn, eidx, slice >> 1};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  double u1 = u53(w[0], w[1]);
  double u2 = u53(w[2], w[3]);
  const double a = ORC_PI * (2.0 * u2);
  return sqrt(-2.0 * log(u1)) * ((slice & 1) ? sin(a) : cos(a));
}

/* ---- longitudinal resonator wake ----------------------------------------- */
double orc_wake_w(double rs, double q, double fr, double tau) {
  double wr = ORC_TWOPI * fr;
  double alpha = wr / (2 * q);
  if (tau == 0) return alpha * rs;
  double wb = sqrt(wr * wr - alpha * alpha);
  return 2 * alpha * rs * exp(-alpha * tau) * (cos(wb * tau) - alpha / wb * sin(wb * tau));
}

/* Binned wake kick over the whole beam: nearest-bin integer histogram of z
 * over [zmin, zmax), potential at each bin from the bins ahead of it plus
 * half its own field (W(0) = alpha*Rs), delta -= V/E0. */
static void orc_longwake(const accp_elem *e, double *r, int n) {
  if (!e->on || n <= 0) return;
  int nb = e->nbins;
  double dz = (e->wake_zmax - e->wake_zmin) / nb;
  double *cnt = (double *)calloc((size_t)nb, sizeof(double));
  double *wt = (double *)malloc((size_t)nb * sizeof(double));
  double *v = (double *)malloc((size_t)nb * sizeof(double));
  int *bin = (int *)malloc((size_t)n * sizeof(int));
  for (int i = 0; i < n; i++) {
    double *p = r + 6 * i;
    bin[i] = -1;
    if (isnan(p[0])) continue;
    double f = floor((p[4] - e->wake_zmin) / dz);
    if (f >= 0 && f < nb) {
      bin[i] = (int)f;
      cnt[bin[i]] += 1.0;
    }
  }
  for (int d = 0; d < nb; d++) wt[d] = orc_wake_w(e->wake_rs, e->wake_q, e->wake_fr, d * dz / ORC_C0);
  double qm = e->wake_charge / n;
  for (int k = 0; k < nb; k++) {
    double s = 0;
    for (int m = k; m < nb; m++) s += cnt[m] * wt[m - k];
    v[k] = qm * s;
  }
  double e0 = e->energy * 1e6;
  for (int i = 0; i < n; i++)
    if (bin[i] >= 0) r[6 * i + 5] -= v[bin[i]] / e0;
  free(cnt); free(wt); free(v); free(bin);
}

void orc_wake_pairwise(const accp_elem *e, const double *r, int n, double *ddelta) {
  double qm = e->wake_charge / n, e0 = e->energy * 1e6;
  for (int i = 0; i < n; i++) {
    ddelta[i] = 0;
    const double *pi = r + 6 * i;
    if (isnan(pi[0])) continue;
    double s = 0;
    for (int j = 0; j < n; j++) {
      const double *pj = r + 6 * j;
      if (isnan(pj[0])) continue;
      double dz = pj[4] - pi[4];
      if (j == i) s += orc_wake_w(e->wake_rs, e->wake_q, e->wake_fr, 0);
      else if (dz > 0) s += orc_wake_w(e->wake_rs, e->wake_q, e->wake_fr, dz / ORC_C0);
    }
    ddelta[i] = -qm * s / e0;
  }
}

/* ---- elements ------------------------------------------------------------ */
static int loss_code(const double *p, const accp_elem *e) {
  return (p[0] <= e->xlim[0]) * 1 + (p[0] >= e->xlim[1]) * 2 + (p[2] <= e->ylim[0]) * 3 +
         (p[2] >= e->ylim[1]) * 4;
}
static int outside(const double *p, const accp_elem *e) {
  return p[0] <= e->xlim[0] || p[0] >= e->xlim[1] || p[2] <= e->ylim[0] || p[2] >= e->ylim[1];
}

/* Half of the ideal wiggler's classical radiation loss per slice, angles
 * kept (accel.cpp:1495-1504). */
static void wig_rad_half(double *p, const accp_elem *e) {
  const double crad = ORC_CGAMMA * e->energy * e->energy * e->energy / (ORC_TWOPI * 1e9);
  const double h = e->wig_h, dp = p[5];
  const double yp = p[3] / (1 + dp), xp = p[1] / (1 + dp);
  p[5] = p[5] - crad * sq(1 + p[5]) * h * h * e->len / e->nslice / 2.0;
  p[1] = xp * (1 + p[5]);
  p[3] = yp * (1 + p[5]);
}

/* Lower index of the grid cell holding v: bisection over a monotonic grid,
 * clamped to [0, n-2] (Base_interp::locate, interp_1d.h:21-37, with mm = 2;
 * interp2 builds a fresh interpolator per call, so hunt() is never used). */
static int wt_locate(const double *xx, int n, double v) {
  const int ascnd = xx[n - 1] >= xx[0];
  int jl = 0, ju = n - 1;
  while (ju - jl > 1) {
    const int jm = (ju + jl) >> 1;
    if ((v >= xx[jm]) == ascnd) jl = jm;
    else ju = jm;
  }
  return jl < 0 ? 0 : (jl > n - 2 ? n - 2 : jl);
}

/* interp2 (utility.cpp:174-187) through Bilin_interp::interp
 * (interp_2d.h:10-20): tab holds my rows of mx values, tab[iy*mx + ix]. */
static double wt_interp2(const double *tab, const double *xl, int mx, const double *yl, int my, double x,
                         double y) {
  const int i = wt_locate(xl, mx, x), j = wt_locate(yl, my, y);
  const double t = (x - xl[i]) / (xl[i + 1] - xl[i]);
  const double u = (y - yl[j]) / (yl[j + 1] - yl[j]);
  return (1. - t) * (1. - u) * tab[j * mx + i] + t * (1. - u) * tab[j * mx + i + 1] +
         (1. - t) * u * tab[(j + 1) * mx + i] + t * u * tab[(j + 1) * mx + i + 1];
}

/* ATransWake::pass (accel.cpp:1603-1657) with the whole beam as its slice
 * (one thread): particle by particle, a half drift, the pairwise kick from
 * every other live particle that is not behind it, a half drift.  q is the
 * charge per macro-particle in pC from Ne [nC] over all n particles. */
static void orc_transwake(const accp_elem *e, double *r, int n) {
  const double q = e->wake_charge / n * 1000;
  const double kRx = 2 * ORC_PI * e->tw_fr[0] * 1.0e9 / ORC_C0, kRy = 2 * ORC_PI * e->tw_fr[1] * 1.0e9 / ORC_C0;
  const double ax = kRx / 2 / e->tw_q[0], ay = kRy / 2 / e->tw_q[1];
  const double kbx = kRx * sqrt(1.0 - 0.25 / sq(e->tw_q[0])), kby = kRy * sqrt(1.0 - 0.25 / sq(e->tw_q[1]));
  for (int i = 0; i < n; i++) {
    double *p = r + 6 * i;
    if (isnan(p[0])) continue;
    if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
    if (e->on) {
      double dxp = 0, dyp = 0;
      for (int j = 0; j < n; j++) {
        const double *pj = r + 6 * j;
        if (isnan(pj[0]) || i == j) continue;
        const double zij = p[4] - pj[4];
        if (zij > 0) continue;
        dxp = dxp - q * e->tw_w[0] * exp(ax * zij) * sin(kbx * zij) * pj[0] / e->energy / 1.0e6;
        dyp = dyp - q * e->tw_w[1] * exp(ay * zij) * sin(kby * zij) * pj[2] / e->energy / 1.0e6;
      }
      p[1] += dxp * (1 + p[5]);
      p[3] += dyp * (1 + p[5]);
    }
    if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
  }
}

void orc_elem_pass(const accp_elem *e, double *r, int n, const orc_ctx *ctx) {
  if (e->type == ACCP_LONGWAKE) {
    orc_longwake(e, r, n);
    return;
  }
  if (e->type == ACCP_TRANSWAKE) {
    orc_transwake(e, r, n);
    return;
  }
  if (e->type == ACCP_MONITOR) return; /* coordinates untouched (accel.cpp:437-475) */
  /* per-element constants computed once per pass, as the reference does */
  double sl = 0, l1 = 0, l2 = 0, k1 = 0, k2 = 0, irho = 0;
  if (e->type == ACCP_BENDMPOLE || e->type == ACCP_STRMPOLE) {
    sl = e->len / e->nslice;           /* accel.cpp:652-656, :804-808 */
    l1 = sl * ORC_DRIFT1;
    l2 = sl * ORC_DRIFT2;
    k1 = sl * ORC_KICK1;
    k2 = sl * ORC_KICK2;
  }
  if (e->type == ACCP_BENDMPOLE || e->type == ACCP_BEND) irho = e->theta / e->len; /* accel.h:170,200 */
  for (int i = 0; i < n; i++) {
    double *p = r + 6 * i;
    if (isnan(p[0])) continue;
    switch (e->type) {
      case ACCP_DRIFT: /* accel.cpp:158-177 */
        orc_fastdrift(p, e->len / (1.0 + p[5]));
        break;
      case ACCP_QUAD: /* accel.cpp:193-209 */
        orc_quad6(p, e->len, e->k);
        break;
      case ACCP_CORRECTOR: { /* accel.cpp:270-292 */
        double pn = 1 + p[5];
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / pn);
        p[1] += e->kickx / pn;
        p[3] += e->kicky / pn;
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / pn);
        break;
      }
      case ACCP_APERTURE: /* accel.cpp:337-387 */
        if (outside(p, e)) {
          p[4] = loss_code(p, e);
          p[0] = NAN;
          break;
        }
        if (e->len != 0) {
          orc_fastdrift(p, e->len / (1 + p[5]));
          if (outside(p, e)) {
            p[4] = loss_code(p, e);
            p[0] = NAN;
          }
        }
        break;
      case ACCP_BEND: /* accel.cpp:523-550 */
        orc_edge(p, irho, e->e1);
        orc_bend6(p, e->len, e->theta, e->k, 0);
        orc_edge(p, irho, e->e2);
        break;
      case ACCP_BENDMPOLE: { /* accel.cpp:649-705 */
        int mo = e->max_npole - 1;
        /* ABendMPole::set_rad_onoff keeps 0..2 and turns anything else off (accel.h:216-218) */
        const int rad = (e->rad == 1 || e->rad == 2) ? e->rad : 0;
        orc_edge(p, irho, e->e1);
        for (int m = 0; m < e->nslice; m++) {
          if (rad <= 1) {
            orc_hxdrift6(p, l1, irho);
            if (rad) orc_bndthinkickrad(p, e->polyB, k1, irho, e->energy, mo, rad, 0);
            else orc_bndthinkick(p, e->polyB, k1, irho, mo);
            orc_hxdrift6(p, l2, irho);
            if (rad) orc_bndthinkickrad(p, e->polyB, k2, irho, e->energy, mo, rad, 0);
            else orc_bndthinkick(p, e->polyB, k2, irho, mo);
            orc_hxdrift6(p, l2, irho);
            if (rad) orc_bndthinkickrad(p, e->polyB, k1, irho, e->energy, mo, rad, 0);
            else orc_bndthinkick(p, e->polyB, k1, irho, mo);
            orc_hxdrift6(p, l1, irho);
          } else {
            orc_hxdrift6(p, 0.5 * sl, irho);
            double gs = orc_gauss(ctx->seed, ctx->pid0 + (uint64_t)i, ctx->turn,
                                  (uint32_t)ctx->eidx, (uint32_t)m);
            orc_bndthinkickrad(p, e->polyB, sl, irho, e->energy, mo, 2, gs);
            orc_hxdrift6(p, 0.5 * sl, irho);
          }
        }
        orc_edge(p, irho, e->e2);
        break;
      }
      case ACCP_STRMPOLE: /* accel.cpp:800-847 */
        if (e->len == 0) {
          orc_strthinkick(p, e->polyA, e->polyB, 1.0, ACCP_MAX_MPOLE - 1);
        } else {
          for (int m = 0; m < e->nslice; m++) {
            double norm = 1 / (1 + p[5]);
            double nl1 = l1 * norm, nl2 = l2 * norm;
            orc_fastdrift(p, nl1);
            orc_strthinkick(p, e->polyA, e->polyB, k1, ACCP_MAX_MPOLE - 1);
            orc_fastdrift(p, nl2);
            orc_strthinkick(p, e->polyA, e->polyB, k2, ACCP_MAX_MPOLE - 1);
            orc_fastdrift(p, nl2);
            orc_strthinkick(p, e->polyA, e->polyB, k1, ACCP_MAX_MPOLE - 1);
            orc_fastdrift(p, nl1);
          }
        }
        break;
      case ACCP_CAVITY: { /* accel.cpp:903-923 */
        double nv = e->volt / e->energy;
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        if (e->on) p[5] += nv * sin(ORC_TWOPI * e->freq * p[4] / ORC_C0 + e->phis);
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        break;
      }
      case ACCP_MATRIX: { /* accel.cpp:1697-1729 */
        const double *m = e->mat;
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        double tr[6];
        for (int j = 0; j < 6; j++)
          tr[j] = m[j * 6 + 0] * p[0] + m[j * 6 + 1] * p[1] + m[j * 6 + 2] * p[2] + m[j * 6 + 3] * p[3] +
                  m[j * 6 + 4] * p[4] + m[j * 6 + 5] * p[5];
        for (int j = 0; j < 6; j++) p[j] = tr[j];
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        break;
      }
      case ACCP_SYNCRAD: { /* accel.cpp:1813-1854, flags from set_rad_onoff (aline.cpp:273-288) */
        const int damp = e->rad == 3 || e->rad == 4, qe = e->rad == 3 || e->rad == 5;
        const double jz = 4 - e->jx - e->jy; /* accel.h:551 */
        const double alfx = e->u0 / 2. / e->energy * e->jx, alfy = e->u0 / 2. / e->energy * e->jy;
        const double alfz = e->u0 / 2. / e->energy * jz;
        const double sx = sqrt(2. * alfx), sy = sqrt(2. * alfy), sz = sqrt(2. * alfz);
        const double sigx = sqrt(e->emitx * e->betax), sigy = sqrt(e->emity * e->betay);
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        if (damp) {
          p[0] *= (1 - alfx);
          p[2] *= (1 - alfy);
          p[5] *= (1 - alfz);
        }
        if (qe) {
          const uint64_t pid = ctx->pid0 + (uint64_t)i;
          const uint32_t k = (uint32_t)ctx->eidx;
          p[0] += sigx * sx * orc_gauss(ctx->seed, pid, ctx->turn, k, 0);
          p[2] += sigy * sy * orc_gauss(ctx->seed, pid, ctx->turn, k, 1);
          p[5] += e->sigdpp * sz * orc_gauss(ctx->seed, pid, ctx->turn, k, 2);
        }
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        break;
      }
      case ACCP_RFMULT: { /* accel.cpp:1146-1201 */
        const double k = ORC_TWOPI * e->freq / ORC_C0;
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        double om;
        if (e->sigphi > 0)
          om = -k * p[4] + e->phis +
               e->sigphi * orc_gauss(ctx->seed, ctx->pid0 + (uint64_t)i, ctx->turn, (uint32_t)ctx->eidx, 0);
        else
          om = -k * p[4] + e->phis;
        const double rho = sqrt(sq(p[0]) + sq(p[2]));
        const double ph = atan2(p[2], p[0]);
        double AsL = 0, ArhoL = 0, AphiL = 0, nrho = 1.0;
        for (int m = 0; m < e->max_npole; m++) {
          const double nphi = (m + 1) * ph;
          const double td = (e->polyB[m] * cos(nphi) + e->polyA[m] * sin(nphi));
          ArhoL += nrho * td;
          AphiL += nrho * (-e->polyB[m] * sin(nphi) + e->polyA[m] * cos(nphi));
          nrho *= rho;
          AsL += 1.0 / (m + 1) * nrho * td;
        }
        p[1] += (ArhoL * cos(ph) - AphiL * sin(ph)) * cos(om);
        p[3] += (ArhoL * sin(ph) + AphiL * cos(ph)) * cos(om);
        p[5] += k * AsL * sin(om) / (1. + p[5]);
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        break;
      }
      case ACCP_IDEALWIG: { /* accel.cpp:1463-1537 */
        if (e->len == 0) break;
        const double k = 2.0 * ORC_PI / e->wig_lambda, h = e->wig_h;
        const int rad = e->rad == 1 || e->rad == 2; /* AIdealWig::set_rad_onoff (accel.h:448-452) */
        for (int j = 0; j < e->nslice; j++) {
          orc_fastdrift(p, 0.5 * e->len / (1 + p[5]) / e->nslice);
          const double y = p[2], x = p[0];
          if (rad) wig_rad_half(p, e);
          if (e->wig_type == 0) {
            const double y3 = y * y * y, k2 = k * k;
            p[3] -= h * h * (0.5 * y + 1.0 / 3.0 * y3 * k2 + 1.0 / 15.0 * y3 * y * y * k2 * k2) * e->len / e->nslice;
          } else if (e->wig_type == 1) {
            const double x3 = x * x * x, k2 = k * k;
            p[1] -= h * h * (0.5 * x + 1.0 / 3.0 * x3 * k2 + 1.0 / 15.0 * x3 * x * x * k2 * k2) * e->len / e->nslice;
          }
          if (rad) wig_rad_half(p, e);
          orc_fastdrift(p, 0.5 * e->len / (1 + p[5]) / e->nslice);
        }
        break;
      }
      case ACCP_WIGTABLE: { /* accel.cpp:1338-1382 */
        const int mx = e->wt_mx, my = e->wt_my;
        const double *xl = e->wt_table, *yl = xl + mx, *xk = yl + my, *yk = xk + mx * my;
        if (e->len == 0) {
          const double x = p[0], y = p[2];
          p[1] += wt_interp2(xk, xl, mx, yl, my, x, y);
          p[3] += wt_interp2(yk, xl, mx, yl, my, x, y);
        } else {
          for (int j = 0; j < e->nslice; j++) {
            orc_fastdrift(p, 0.5 * e->len / (1 + p[5]) / e->nslice);
            const double x = p[0], y = p[2];
            p[1] += wt_interp2(xk, xl, mx, yl, my, x, y) / e->nslice;
            p[3] += wt_interp2(yk, xl, mx, yl, my, x, y) / e->nslice;
            orc_fastdrift(p, 0.5 * e->len / (1 + p[5]) / e->nslice);
          }
        }
        break;
      }
      case ACCP_CRABCAV: { /* accel.cpp:968-999 */
        const double nvx = e->devoltx / e->energy, nvy = e->devolty / e->energy;
        const double k = ORC_TWOPI * e->freq / ORC_C0;
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        if (e->on) {
          const uint64_t pid = ctx->pid0 + (uint64_t)i;
          const uint32_t ei = (uint32_t)ctx->eidx;
          /* normals in the reference's draw order (accel.cpp:1015-1022):
           * the phase's first when it draws, then x's and y's amplitudes */
          double om, nvxa = nvx, nvya = nvy;
          uint32_t j = 0;
          if (e->sigphi > 0) om = -k * p[4] + e->phis + e->sigphi * orc_gauss(ctx->seed, pid, ctx->turn, ei, j++);
          else om = -k * p[4] + e->phis;
          if (e->sigdvv > 0) {
            nvxa = nvx * (1.0 + e->sigdvv * orc_gauss(ctx->seed, pid, ctx->turn, ei, j));
            nvya = nvy * (1.0 + e->sigdvv * orc_gauss(ctx->seed, pid, ctx->turn, ei, j + 1));
          }
          p[1] += nvxa * sin(om);
          p[3] += nvya * sin(om);
          p[5] += -(nvya * p[2] + nvxa * p[0]) * k * cos(om) / (1 + p[5]);
        }
        if (e->len != 0) orc_fastdrift(p, 0.5 * e->len / (1 + p[5]));
        break;
      }
      default:
        break;
    }
  }
}

/* aline.cpp:331-363 (a Monitor at element 0 leaves coordinates alone, so
 * pass(r,N) is elements 0..n-1), element-major over the beam. */
void orc_line_pass(const accp_elem *elems, int nelem, double *r, int n, int st, int ed,
                   int nturn, uint64_t seed, uint64_t turn0, uint64_t pid0) {
  if (ed < 0 || ed > nelem) ed = nelem;
  orc_ctx ctx = {seed, turn0, pid0, 0};
  for (int t = 0; t < nturn; t++) {
    ctx.turn = turn0 + (uint64_t)t;
    for (int k = st; k < ed; k++) {
      ctx.eidx = k;
      orc_elem_pass(&elems[k], r, n, &ctx);
    }
  }
}

/* aline.cpp:313-329: per turn, one parallel region; thread t owns particles
 * [t*N/T, (t+1)*N/T) and pushes them through every element. */
int orc_line_pass_omp(const accp_elem *elems, int nelem, double *r, int n, int nturn,
                      int nthreads, uint64_t seed, uint64_t turn0, uint64_t pid0) {
  for (int k = 0; k < nelem; k++)
    if ((elems[k].type == ACCP_LONGWAKE || elems[k].type == ACCP_TRANSWAKE) && elems[k].on) return -1;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  omp_set_num_threads(nthreads);
#endif
  for (int t = 0; t < nturn; t++) {
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
      int tid = 0, nth = 1;
#ifdef _OPENMP
      tid = omp_get_thread_num();
      nth = omp_get_num_threads();
#endif
      int s = (int)((long long)tid * n / nth);
      int e = (int)((long long)(tid + 1) * n / nth);
      orc_ctx ctx = {seed, turn0 + (uint64_t)t, pid0 + (uint64_t)s, 0};
      for (int k = 0; k < nelem; k++) {
        ctx.eidx = k;
        orc_elem_pass(&elems[k], r + 6 * (size_t)s, e - s, &ctx);
      }
    }
  }
  return 0;
}

/* ---- optics by tracking --------------------------------------------------- */
/* aline.cpp:900-927 -- 12 rays at +-delta on each axis, central difference,
 * elements [st, ed) called directly (serial). */
void orc_getm66(const accp_elem *elems, int nelem, int st, int ed, double dpp,
                double delta, double m66[36]) {
  double r0[72];
  memset(r0, 0, sizeof(r0));
  for (int j = 0; j < 6; j++) {
    r0[j * 6 + j] = delta;
    r0[36 + j * 6 + j] = -delta;
  }
  for (int i = 0; i < 12; i++) r0[i * 6 + 5] += dpp;
  if (ed < 0 || ed > nelem) ed = nelem;
  orc_ctx ctx = {0, 0, 0, 0};
  for (int k = st; k < ed; k++) {
    ctx.eidx = k;
    orc_elem_pass(&elems[k], r0, 12, &ctx);
  }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) m66[i * 6 + j] = (r0[j * 6 + i] - r0[j * 6 + i + 36]) / 2.0 / delta;
}

/* aline.cpp:1032-1043 */
void orc_calctune(const accp_elem *elems, int nelem, double dpp, double delta, double *tunex,
                  double *tuney) {
  double m[36];
  orc_getm66(elems, nelem, 0, nelem, dpp, delta, m);
  *tunex = acos((m[0] + m[7]) * 0.5) / 2.0 / ORC_PI;
  *tuney = acos((m[14] + m[21]) * 0.5) / 2.0 / ORC_PI;
}

/* aline.cpp:1108-1126 */
void orc_calcchrom(const accp_elem *elems, int nelem, double dp, double delta, double *chromx,
                   double *chromy) {
  double dpp = 1.0e-5, m[36];
  orc_getm66(elems, nelem, 0, nelem, dp - dpp, delta, m);
  double tx0 = acos((m[0] + m[7]) * 0.5) / 2.0 / ORC_PI;
  double ty0 = acos((m[14] + m[21]) * 0.5) / 2.0 / ORC_PI;
  orc_getm66(elems, nelem, 0, nelem, dp + dpp, delta, m);
  double tx = acos((m[0] + m[7]) * 0.5) / 2.0 / ORC_PI;
  double ty = acos((m[14] + m[21]) * 0.5) / 2.0 / ORC_PI;
  *chromx = (tx - tx0) / dpp / 2.0;
  *chromy = (ty - ty0) / dpp / 2.0;
}

/* Field a tune/chromaticity knob moves: Quad K, StrMPole PolyB[1] (tune) or
 * StrMPole PolyB[2] (chromaticity); aline.cpp:1201-1213, :1324-1330. */
static double *knob_field(accp_elem *e, int kind) {
  if (kind == ORC_KNOB_TUNE) {
    if (e->type == ACCP_QUAD) return &e->k;
    if (e->type == ACCP_STRMPOLE) return &e->polyB[1];
  } else if (e->type == ACCP_STRMPOLE) {
    return &e->polyB[2];
  }
  return NULL;
}

static void measure(const accp_elem *elems, int nelem, int kind, double *a, double *b) {
  if (kind == ORC_KNOB_TUNE) orc_calctune(elems, nelem, 0, ORC_DELTA_DIFF, a, b);
  else orc_calcchrom(elems, nelem, 0, ORC_DELTA_DIFF, a, b);
}

/* aline.cpp:1187-1242 (tune) and :1310-1355 (chromaticity): forward
 * difference with +1e-5 on every member of a family, values restored. */
void orc_calcresp(accp_elem *elems, int nelem, int kind, const int *fam1, int n1,
                  const int *fam2, int n2, double resp[4]) {
  double a0, b0;
  measure(elems, nelem, kind, &a0, &b0);
  const double delta = 1.0e-5;
  const int *fams[2] = {fam1, fam2};
  int nf[2] = {n1, n2};
  for (int kn = 0; kn < 2; kn++) {
    double *old = (double *)malloc(sizeof(double) * (size_t)(nf[kn] > 0 ? nf[kn] : 1));
    for (int j = 0; j < nf[kn]; j++) {
      double *f = knob_field(&elems[fams[kn][j]], kind);
      if (!f) continue;
      old[j] = *f;
      *f = old[j] + delta;
    }
    double a, b;
    measure(elems, nelem, kind, &a, &b);
    resp[kn * 2 + 0] = (a - a0) / delta;
    resp[kn * 2 + 1] = (b - b0) / delta;
    for (int j = 0; j < nf[kn]; j++) {
      double *f = knob_field(&elems[fams[kn][j]], kind);
      if (f) *f = old[j];
    }
    free(old);
  }
}

/* aline.cpp:1243-1309 and :1356-1418 -- one 2x2 Newton step. */
void orc_correct(accp_elem *elems, int nelem, int kind, const int *fam1, int n1,
                 const int *fam2, int n2, double *tx, double *ty) {
  double a0, b0, resp[4];
  measure(elems, nelem, kind, &a0, &b0);
  orc_calcresp(elems, nelem, kind, fam1, n1, fam2, n2, resp);
  double detr = resp[0] * resp[3] - resp[1] * resp[2];
  double d1 = -(resp[3] * (a0 - *tx) - resp[2] * (b0 - *ty)) / detr;
  double d2 = -(-resp[1] * (a0 - *tx) + resp[0] * (b0 - *ty)) / detr;
  for (int j = 0; j < n1; j++) {
    double *f = knob_field(&elems[fam1[j]], kind);
    if (f) *f += d1;
  }
  for (int j = 0; j < n2; j++) {
    double *f = knob_field(&elems[fam2[j]], kind);
    if (f) *f += d2;
  }
  measure(elems, nelem, kind, tx, ty);
}

/* aline.cpp:465-501 */
double orc_setrfphase(accp_elem *elems, int nelem) {
  double r[6] = {0, 0, 0, 0, 0, 0};
  /* LongWake (no reference counterpart) is off for the probe, which would
   * otherwise see the whole bunch charge's own field; ATransWake, the
   * reference's only wake, cannot kick a lone particle (accel.cpp:1638). */
  int *wake_on = (int *)malloc(sizeof(int) * (size_t)(nelem > 0 ? nelem : 1));
  for (int k = 0; k < nelem; k++) {
    if (elems[k].type == ACCP_CAVITY) elems[k].on = 0;
    if (elems[k].type == ACCP_BENDMPOLE) elems[k].rad = 1;
    wake_on[k] = elems[k].on;
    if (elems[k].type == ACCP_LONGWAKE) elems[k].on = 0;
  }
  orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
  double delta = -r[5];
  for (int k = 0; k < nelem; k++) {
    if (elems[k].type == ACCP_BENDMPOLE) elems[k].rad = 0;
    if (elems[k].type == ACCP_LONGWAKE) elems[k].on = wake_on[k];
  }
  free(wake_on);
  double volt = 0, energy = 0;
  int ncav = 0;
  for (int k = 0; k < nelem; k++)
    if (elems[k].type == ACCP_CAVITY) {
      volt += elems[k].volt;
      energy = elems[k].energy;
      ncav++;
    }
  if (!ncav) return 0;
  double phis = asin(delta * energy / volt);
  for (int k = 0; k < nelem; k++)
    if (elems[k].type == ACCP_CAVITY) elems[k].phis = phis;
  return phis;
}

/* aline.cpp:607-715 (the loss-turn branch, :644-666, only prints). */
double orc_trackda(const accp_elem *elems, int nelem, double xmax, double xmin, double ymax,
                   int nturn, int nray, int nstep, double *rx, double *ry, int negweight) {
  if (nray < 3) nray = 3;
  if (nray > 51) nray = 51;
  if (nstep < 2) nstep = 2;
  double dx = xmax - xmin;
  if (dx < 0) return -1;
  int np = nray * nstep;
  double *xl = (double *)malloc(sizeof(double) * (size_t)np);
  double *yl = (double *)malloc(sizeof(double) * (size_t)np);
  double *r = (double *)calloc((size_t)np * 6, sizeof(double));
  for (int i = 0; i < nray; i++)
    for (int j = 0; j < nstep; j++) {
      int q = i * nstep + j;
      xl[q] = (xmin + dx * j / (nstep - 1)) * cos(i * ORC_PI / (nray - 1));
      yl[q] = (xmin + dx * j / (nstep - 1)) * sin(i * ORC_PI / (nray - 1)) * ymax / xmax;
      r[q * 6 + 0] = xl[q];
      r[q * 6 + 2] = yl[q];
      r[q * 6 + 4] = 1.0e-5;
    }
  orc_line_pass(elems, nelem, r, np, 0, nelem, nturn, 0, 0, 0);
  for (int i = 0; i < nray; i++) {
    rx[i] = xl[i * nstep + nstep - 1];
    ry[i] = yl[i * nstep + nstep - 1];
    for (int j = 0; j < nstep; j++)
      if (isnan(r[(i * nstep + j) * 6])) {
        rx[i] = xl[i * nstep + j];
        ry[i] = yl[i * nstep + j];
        break;
      }
  }
  double s = 0;
  for (int i = 1; i < nray; i++) {
    double rd1 = sqrt(rx[i - 1] * rx[i - 1] + ry[i - 1] * ry[i - 1]);
    double rd2 = sqrt(rx[i] * rx[i] + ry[i] * ry[i]);
    double dphi = atan2(ry[i], rx[i]) - atan2(ry[i - 1], rx[i - 1]);
    if (rx[i] <= 0) s += .5 * rd1 * rd2 * sin(dphi);
    else if (negweight) s += .25 * rd1 * rd2 * sin(dphi);
    else s += .5 * rd1 * rd2 * sin(dphi);
  }
  free(xl); free(yl); free(r);
  return s;
}

/* Solve A x = b (n <= 6) by Gaussian elimination with partial pivoting. */
static void solve_small(int n, double *A, double *b) {
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(A[r * n + c]) > fabs(A[piv * n + c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < n; k++) {
        double t = A[c * n + k];
        A[c * n + k] = A[piv * n + k];
        A[piv * n + k] = t;
      }
      double t = b[c];
      b[c] = b[piv];
      b[piv] = t;
    }
    for (int r = c + 1; r < n; r++) {
      double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double s = b[r];
    for (int k = r + 1; k < n; k++) s -= A[r * n + k] * b[k];
    b[r] = s / A[r * n + r];
  }
}

static void set_flags(accp_elem *elems, int nelem, int rad, int cav) {
  for (int k = 0; k < nelem; k++) {
    if (elems[k].type == ACCP_BENDMPOLE) elems[k].rad = rad;
    if (elems[k].type == ACCP_CAVITY) elems[k].on = cav;
  }
}

/* aline.cpp:838-870 */
void orc_getco4(accp_elem *elems, int nelem, double dp, double rc[4]) {
  set_flags(elems, nelem, 0, 0);
  double m[36], r[6] = {0, 0, 0, 0, 0, dp}, A[16];
  orc_getm66(elems, nelem, 0, nelem, dp, ORC_DELTA_DIFF, m);
  orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) A[i * 4 + j] = (i == j ? 1.0 : 0.0) - m[i * 6 + j];
    rc[i] = r[i];
  }
  solve_small(4, A, rc);
}

/* aline.cpp:871-889 */
void orc_getco6(accp_elem *elems, int nelem, double rc[6]) {
  set_flags(elems, nelem, 1, 1);
  double m[36], r[6] = {0, 0, 0, 0, 0, 0}, A[36];
  orc_getm66(elems, nelem, 0, nelem, 0.0, ORC_DELTA_DIFF, m);
  orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) A[i * 6 + j] = (i == j ? 1.0 : 0.0) - m[i * 6 + j];
    rc[i] = r[i];
  }
  solve_small(6, A, rc);
}

/* aline.cpp:1095-1107 */
void orc_calctwiss0(const accp_elem *elems, int nelem, double dp, double tw[7]) {
  double m[36];
  orc_getm66(elems, nelem, 0, nelem, dp, ORC_DELTA_DIFF, m);
  double psix = acos((m[0] + m[7]) * 0.5), psiy = acos((m[14] + m[21]) * 0.5);
  tw[0] = 0;
  tw[1] = psix;
  tw[2] = psiy;
  tw[3] = m[1] / sin(psix);
  tw[4] = (m[0] - m[7]) / sin(psix) / 2.0;
  tw[5] = m[15] / sin(psiy);
  tw[6] = (m[14] - m[21]) / sin(psiy) / 2.0;
}

static double line_length(const accp_elem *elems, int st, int ed) {
  double s = 0;
  for (int k = st; k < ed; k++) s += elems[k].len;
  return s;
}

/* aline.cpp:1044-1094: Twiss at the start, then carried point to point by
 * the transfer matrices between the listed elements. */
void orc_calctwiss(const accp_elem *elems, int nelem, const int *idx, int n, double dp,
                   double *tw) {
  double t0[7], T[36];
  orc_calctwiss0(elems, nelem, dp, t0);
  t0[1] = t0[2] = 0;
  int id = idx[0];
  orc_getm66(elems, nelem, 0, id, dp, ORC_DELTA_DIFF, T);
  double len01 = line_length(elems, 0, id);
  for (int i = 0; i < n; i++) {
    const double bx = t0[3], ax = t0[4], by = t0[5], ay = t0[6];
    double *t1 = tw + 7 * i;
    t1[0] = t0[0] + len01;
    t1[3] = 1.0 / bx * (sq(T[0] * bx - T[1] * ax) + sq(T[1]));
    t1[5] = 1.0 / by * (sq(T[14] * by - T[15] * ay) + sq(T[15]));
    t1[4] = -1.0 / bx * ((T[0] * bx - T[1] * ax) * (T[6] * bx - T[7] * ax) + T[1] * T[7]);
    t1[6] = -1.0 / by * ((T[14] * by - T[15] * ay) * (T[20] * by - T[21] * ay) + T[15] * T[21]);
    t1[1] = t0[1] + atan2(T[1], T[0] * bx - T[1] * ax);
    t1[2] = t0[2] + atan2(T[15], T[14] * by - T[15] * ay);
    for (int k = 0; k < 7; k++) t0[k] = t1[k];
    if (i < n - 1) {
      int id2 = idx[i + 1];
      orc_getm66(elems, nelem, id, id2, dp, ORC_DELTA_DIFF, T);
      len01 = line_length(elems, id, id2);
      id = id2;
    }
  }
}

/* Least-squares slope of y on x, centred on the mean of x. */
static double lsq_slope(int n, const double *x, const double *y) {
  double sx = 0;
  for (int i = 0; i < n; i++) sx += x[i];
  const double mx = sx / n;
  double stt = 0, sty = 0;
  for (int i = 0; i < n; i++) {
    const double t = x[i] - mx;
    stt += t * t;
    sty += t * y[i];
  }
  return sty / stt;
}

/* aline.cpp:1127-1158 */
void orc_calcchrom2(const accp_elem *elems, int nelem, double c[4]) {
  double cx[11], cy[11], dp[11];
  for (int i = 0; i < 11; i++) {
    dp[i] = 1.0e-3 * (i - 6.0);
    orc_calcchrom(elems, nelem, dp[i], ORC_DELTA_DIFF, &cx[i], &cy[i]);
  }
  c[2] = lsq_slope(11, dp, cx) / 2.0;
  c[3] = lsq_slope(11, dp, cy) / 2.0;
  orc_calcchrom(elems, nelem, 0.0, ORC_DELTA_DIFF, &c[0], &c[1]);
}

/* aline.cpp:781-836; resetfirstelement (aline.cpp:557-606) as a rotated copy */
double orc_trackma(const accp_elem *elems, int nelem, double dppmax, double dppmin, int nturn,
                   int nstep, const int *locs, int nloc, double *dpp_plus, double *dpp_minus) {
  const int np = 2 * (nstep + 1);
  double *r = (double *)malloc(sizeof(double) * 6 * (size_t)np);
  accp_elem *rot = (accp_elem *)malloc(sizeof(accp_elem) * (size_t)nelem);
  const double stepsize = (dppmax - dppmin) / nstep;
  double avg = 0;
  for (int i = 0; i < nloc; i++) {
    const int e0 = locs[i];
    for (int k = 0; k < nelem; k++) rot[k] = elems[(e0 + k) % nelem];
    memset(r, 0, sizeof(double) * 6 * (size_t)np);
    for (int j = 0; j < nstep + 1; j++) {
      r[j * 12 + 5] = dppmin + stepsize * j;
      r[j * 12 + 6 + 5] = -(dppmin + stepsize * j);
    }
    orc_line_pass(rot, nelem, r, np, 0, nelem, nturn, 0, 0, 0);
    dpp_plus[i] = dppmin;
    dpp_minus[i] = -dppmin;
    for (int j = 1; j < nstep + 1; j++) {
      if (!isnan(r[j * 12])) dpp_plus[i] = dppmin + stepsize * j;
      if (!isnan(r[j * 12 + 6])) dpp_minus[i] = -(dppmin + stepsize * j);
    }
    avg += fmin(dpp_plus[i], -dpp_minus[i]);
  }
  free(r);
  free(rot);
  return avg / nloc;
}

/* utility.cpp:322-373 */
int orc_beam_stat(const double *r, int n, double turn, double *stat) {
  double ar[6] = {0}, sg[6] = {0}, cr[6] = {0};
  int nlive = 0;
  for (int j = 0; j < n; j++) {
    const double *p = r + 6 * j;
    if (isnan(p[0])) continue;
    nlive++;
    for (int i = 0; i < 6; i++) ar[i] += p[i];
  }
  for (int i = 0; i < 6; i++) ar[i] /= nlive;
  for (int j = 0; j < n; j++) {
    const double *p = r + 6 * j;
    if (isnan(p[0])) continue;
    for (int i = 0; i < 6; i++) sg[i] += sq(p[i] - ar[i]);
  }
  for (int i = 0; i < 6; i++) {
    sg[i] /= nlive;
    sg[i] = sqrt(sg[i]);
  }
  static const int pa[6] = {0, 2, 0, 1, 2, 3}, pb[6] = {1, 3, 5, 5, 4, 4};
  for (int j = 0; j < n; j++) {
    const double *p = r + 6 * j;
    if (isnan(p[0])) continue;
    for (int i = 0; i < 6; i++) cr[i] += (p[pa[i]] - ar[pa[i]]) * (p[pb[i]] - ar[pb[i]]);
  }
  for (int i = 0; i < 6; i++) cr[i] /= nlive;
  stat[0] = turn;
  stat[1] = nlive;
  for (int i = 0; i < 6; i++) {
    stat[2 + i] = ar[i];
    stat[8 + i] = sg[i];
    stat[14 + i] = cr[i];
  }
  return nlive;
}

/* ---- tune diagnostics: NAFF (utility.cpp:21-96) ---------------------------
 * The spectrum only locates the strongest line; any correct DFT gives the
 * same power up to round-off, so the restatement uses its own iterative
 * radix-2 transform (Cooley-Tukey, complex, imaginary part zero) where the
 * reference calls realft.  The refinement is the golden-section search of
 * the windowed projection, with the reference's constants and steps: its
 * result depends on them bit for bit. */
static void orc_fft(double *re, double *im, int n) {
  for (int i = 1, j = 0; i < n; i++) { /* bit-reversal permutation */
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int h = len >> 1;
    for (int k = 0; k < h; k++) {
      const double a = -2 * ORC_PI * k / len, wr = cos(a), wi = sin(a);
      for (int s = 0; s < n; s += len) {
        const int p = s + k, q = p + h;
        const double tr = re[q] * wr - im[q] * wi, ti = re[q] * wi + im[q] * wr;
        re[q] = re[p] - tr;
        im[q] = im[p] - ti;
        re[p] += tr;
        im[p] += ti;
      }
    }
  }
}

/* fprojection (utility.cpp:21-41): minus the modulus of the projection of
 * the signal, windowed by sin^4(pi i / N), on exp(-2 pi i nv j). */
double orc_fprojection(double nv, int n, const double *x) {
  double ry = 0, iy = 0;
  for (int i = 0; i < n; i++) {
    const double rz = cos(2 * ORC_PI * nv * i), iz = -sin(2 * ORC_PI * nv * i);
    const double w = pow(sin(ORC_PI * i / n / 1.0), 4);
    ry += rz * x[i] * w;
    iy += iz * x[i] * w;
  }
  return -sqrt(ry * ry + iy * iy);
}

/* Golden-section search (Kiefer 1953) for a minimum of fprojection in the
 * bracket (ax, bx, cx), as mins.h:72-110 runs it: section ratio 0.61803399,
 * the new point placed in the larger part, stop when the bracket is below
 * 3e-8 (|x1| + |x2|). */
static double orc_golden(double ax, double bx, double cx, int n, const double *x) {
  const double R = 0.61803399, C = 1.0 - R, tol = 3.0e-8;
  double x0 = ax, x3 = cx, x1, x2;
  if (fabs(cx - bx) > fabs(bx - ax)) {
    x1 = bx;
    x2 = bx + C * (cx - bx);
  } else {
    x2 = bx;
    x1 = bx - C * (bx - ax);
  }
  double f1 = orc_fprojection(x1, n, x), f2 = orc_fprojection(x2, n, x);
  while (fabs(x3 - x0) > tol * (fabs(x1) + fabs(x2))) {
    if (f2 < f1) {
      x0 = x1;
      x1 = x2;
      x2 = R * x2 + C * x3;
      f1 = f2;
      f2 = orc_fprojection(x2, n, x);
    } else {
      x3 = x2;
      x2 = x1;
      x1 = R * x1 + C * x0;
      f2 = f1;
      f1 = orc_fprojection(x1, n, x);
    }
  }
  return f1 < f2 ? x1 : x2;
}

/* naff (utility.cpp:48-96): the strongest line of the spectrum between xlow
 * and xhi (at least 0.9/N, at most 1/2) brackets the golden-section search
 * to one bin either side.  The spectrum is read as the reference's packed
 * real transform: pair 0 holds F_0 and F_N/2, pair k > 0 holds F_k. */
double orc_naff(int n, const double *data, double xlow, double xhi) {
  double *re = (double *)malloc(sizeof(double) * (size_t)n), *im = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    re[i] = data[i];
    im[i] = 0;
  }
  orc_fft(re, im, n);
  xlow = xlow > 0.9 / n ? xlow : 0.9 / n;
  xhi = 0.5 < xhi ? 0.5 : xhi;
  const unsigned ilow = ((int)(xlow * n * 2) >> 1) << 1, ihi = ((int)(xhi * n * 2) >> 1) << 1;
#define ORC_PAIR_POWER(i) \
  ((i) == 0 ? sq(re[0]) + sq(re[n / 2]) : sq(re[(i) / 2]) + sq(im[(i) / 2]))
  double imx = ilow, fm = ORC_PAIR_POWER(ilow);
  for (unsigned i = ilow + 2; i < ihi; i += 2) {
    const double fnew = ORC_PAIR_POWER(i);
    if (fnew > fm) {
      imx = i;
      fm = fnew;
    }
  }
#undef ORC_PAIR_POWER
  free(re);
  free(im);
  const double lo = (imx - 2) * 1.0 / n / 2.0, hi = (imx + 2) * 1.0 / n / 2.0;
  xlow = xlow > lo ? xlow : lo;
  xhi = xhi < hi ? xhi : hi;
  return orc_golden(xlow, (xlow + xhi) / 2.0, xhi, n, data);
}

static unsigned orc_pow2_at_least(unsigned nturn) {
  unsigned nt = 1;
  while (nt < nturn) nt <<= 1;
  return nt;
}

/* ALine::tracktune (aline.cpp:1159-1186): one particle from (x0, 0, y0, 0,
 * 0, 0), nt - 1 turns (nt the power of two >= nturn), NAFF of the nt
 * positions (the start included). */
void orc_tracktune(const accp_elem *elems, int nelem, unsigned nturn, double x0, double y0, double xlow,
                   double xhi, double ylow, double yhi, double tune[2]) {
  const unsigned nt = orc_pow2_at_least(nturn);
  double *xl = (double *)malloc(sizeof(double) * nt), *yl = (double *)malloc(sizeof(double) * nt);
  double r[6] = {x0, 0, y0, 0, 0, 0};
  xl[0] = r[0];
  yl[0] = r[2];
  for (unsigned i = 1; i < nt; i++) {
    orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
    xl[i] = r[0];
    yl[i] = r[2];
  }
  tune[0] = orc_naff((int)nt, xl, xlow, xhi);
  tune[1] = orc_naff((int)nt, yl, ylow, yhi);
  free(xl);
  free(yl);
}

/* ALine::trackDiffusion (aline.cpp:1421-1484): per particle the tunes of
 * the two halves of nt - 1 turns; dif = max(-4 ln nt, ln(2 |dnu| / nt));
 * lost particles give NaN and -3 to the returned sum. */
double orc_trackdiffusion(const accp_elem *elems, int nelem, int nturn, int np, const double *x0,
                          const double *y0, double *nux, double *nuy, double *dif, double xlow, double xhi,
                          double ylow, double yhi) {
  const unsigned nt = orc_pow2_at_least((unsigned)nturn);
  const double mdiffu = -4 * log(nt * 1.0);
  double *xl = (double *)malloc(sizeof(double) * nt), *yl = (double *)malloc(sizeof(double) * nt);
  double tdif = 0;
  for (int ip = 0; ip < np; ip++) {
    double r[6] = {x0[ip], 0, y0[ip], 0, 0, 0};
    xl[0] = r[0];
    yl[0] = r[2];
    for (unsigned j = 1; j < nt; j++) {
      orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
      xl[j] = r[0];
      yl[j] = r[2];
    }
    if (isnan(r[0])) {
      nux[ip] = nuy[ip] = dif[ip] = NAN;
      tdif += -3.0;
    } else {
      const int h = (int)(nt >> 1);
      const double nux0 = orc_naff(h, xl, xlow, xhi), nuy0 = orc_naff(h, yl, ylow, yhi);
      const double nux1 = orc_naff(h, xl + h, xlow, xhi), nuy1 = orc_naff(h, yl + h, ylow, yhi);
      nux[ip] = nux0;
      nuy[ip] = nuy0;
      const double d = log(sqrt(sq(nux1 - nux0) + sq(nuy1 - nuy0)) * 2.0 / nt);
      dif[ip] = mdiffu > d ? mdiffu : d;
      tdif += dif[ip];
    }
  }
  free(xl);
  free(yl);
  return tdif;
}

/* Least-squares slope of y on x (the textbook centred form that Fitab
 * evaluates, fitab.h:37-57). */
static double orc_slope(int n, const double *x, const double *y) {
  double sx = 0, sy = 0, st2 = 0, b = 0;
  for (int i = 0; i < n; i++) {
    sx += x[i];
    sy += y[i];
  }
  const double sxoss = sx / n;
  for (int i = 0; i < n; i++) {
    const double t = x[i] - sxoss;
    st2 += t * t;
    b += t * y[i];
  }
  (void)sy;
  return b / st2;
}

/* ALine::trackdnudexy (aline.cpp:716-779): tune shift with amplitude.  Six
 * particles at x = 1e-5 + k xm / 5 (y = 1e-5 + 0.02 x), nt turns each, tunes
 * by NAFF within 0.1 of the linear tunes; slopes of nux and nuy against
 * x^2 / betx; then the same in y.  d = {dnux/dex, dnuy/dex, dnux/dey,
 * dnuy/dey}. */
void orc_trackdnudexy(const accp_elem *elems, int nelem, unsigned nturn, double xm, double ym, double d[4]) {
  double tw[7];
  orc_calctwiss0(elems, nelem, 0.0, tw);
  const double tunex = tw[1] / 2 / ORC_PI, tuney = tw[2] / 2 / ORC_PI;
  const unsigned np = 6, nt = orc_pow2_at_least(nturn);
  const double tiny = 1.0e-5;
  double xt[6], yt[6], de[6];
  double *xl = (double *)malloc(sizeof(double) * nt), *yl = (double *)malloc(sizeof(double) * nt);
  const double xlo = 0 > tunex - 0.1 ? 0 : tunex - 0.1, xhi = 0.5 < tunex + 0.1 ? 0.5 : tunex + 0.1;
  const double ylo = 0 > tuney - 0.1 ? 0 : tuney - 0.1, yhi = 0.5 < tuney + 0.1 ? 0.5 : tuney + 0.1;
  for (int plane = 0; plane < 2; plane++) {
    for (unsigned k = 0; k < np; k++) {
      double r[6] = {0, 0, 0, 0, 0, 0};
      if (plane == 0) {
        r[0] = tiny + k * 1.0 / (np - 1) * xm;
        r[2] = tiny + r[0] * .02;
        de[k] = r[0] * r[0] / tw[3];
      } else {
        r[2] = tiny + k * 1.0 / (np - 1) * ym;
        r[0] = tiny + r[2] * .02;
        de[k] = r[2] * r[2] / tw[5];
      }
      for (unsigned i = 0; i < nt; i++) {
        xl[i] = r[0];
        yl[i] = r[2];
        orc_line_pass(elems, nelem, r, 1, 0, nelem, 1, 0, 0, 0);
      }
      xt[k] = orc_naff((int)nt, xl, xlo, xhi);
      yt[k] = orc_naff((int)nt, yl, ylo, yhi);
    }
    d[2 * plane] = orc_slope((int)np, de, xt);
    d[2 * plane + 1] = orc_slope((int)np, de, yt);
  }
  free(xl);
  free(yl);
}
