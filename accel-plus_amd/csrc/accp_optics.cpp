// accp_optics.cpp -- ALine's lattice services (transfer matrix, tunes,
// chromaticities, their responses and corrections, RF phase, dynamic
// aperture) on top of the GPU tracker.
//
// Host code over the public entry points of accp.h: rays and particles are
// tracked by accp_track, family strengths and flags are changed in the host
// element records and re-uploaded with accp_lattice_update.  Each function
// keeps the reference's arithmetic (operation order, constants) so that its
// results follow the reference's up to the tracking tolerance; built with
// -ffp-contract=off like the rest of the host code.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <exception>
#include <string>
#include <utility>
#include <vector>

#include "../../include/accp.h"
#include "../host/turnmap.hpp"
#include "accp_error.h"

namespace {

constexpr double PI = 3.1415926535897932384626;  // accel.h:23
constexpr double DPP_CHROM = 1.0e-5;             // aline.cpp:1110
constexpr double DK_RESP = 1.0e-5;               // aline.cpp:1199, :1322

int fail(int code, const std::string &msg) { return accp::set_error(code, msg.c_str()); }

struct Records {
  std::vector<accp_elem> e;
  int load(accp_lattice *L) {
    int n = 0;
    int rc = accp_lattice_nelem(L, &n);
    if (rc) return rc;
    e.resize(n);
    return accp_lattice_get(L, e.data(), n);
  }
  int store(accp_lattice *L) { return accp_lattice_update(L, e.data(), (int)e.size()); }
};

// The 12 rays of getm66 (aline.cpp:911-914): +delta then -delta on each axis,
// all shifted by dp in delta.
void m66_rays(double dp, double delta, double *r0) {
  for (int i = 0; i < 72; ++i) r0[i] = 0.0;
  for (int j = 0; j < 6; ++j) {
    r0[j * 6 + j] = delta;
    r0[36 + j * 6 + j] = -delta;
  }
  for (int i = 0; i < 12; ++i) r0[i * 6 + 5] += dp;
}

// aline.cpp:920-925
void m66_from_rays(const double *r, double delta, double *m66) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) m66[i * 6 + j] = (r[j * 6 + i] - r[j * 6 + i + 36]) / 2.0 / delta;
}

// aline.cpp:1040-1041
void tunes_from_m66(const double *m, double *tx, double *ty) {
  *tx = std::acos((m[0] + m[7]) * 0.5) / 2.0 / PI;
  *ty = std::acos((m[14] + m[21]) * 0.5) / 2.0 / PI;
}

int measure(accp_lattice *L, int quantity, double *a, double *b) {
  double v[2];
  int rc = quantity == ACCP_TUNE ? accp_calctune(L, 0.0, ACCP_DELTA_DIFF, v)
                                 : accp_calcchrom(L, 0.0, ACCP_DELTA_DIFF, v);
  *a = v[0];
  *b = v[1];
  return rc;
}

// The field a tune or chromaticity knob moves (aline.cpp:1201-1213,
// :1324-1330): Quad K or StrMPole PolyB[1] for tunes, StrMPole PolyB[2] for
// chromaticities.
double *knob_field(accp_elem &e, int quantity) {
  if (quantity == ACCP_TUNE) {
    if (e.type == ACCP_QUAD) return &e.k;
    if (e.type == ACCP_STRMPOLE) return &e.polyB[1];
  } else if (e.type == ACCP_STRMPOLE) {
    return &e.polyB[2];
  }
  return nullptr;
}

int check_families(const Records &R, int quantity, const int32_t *fam[2], const int n[2]) {
  if (quantity != ACCP_TUNE && quantity != ACCP_CHROM) return fail(ACCP_E_ARG, "quantity must be ACCP_TUNE or ACCP_CHROM");
  for (int k = 0; k < 2; ++k) {
    if (n[k] < 1 || !fam[k]) return fail(ACCP_E_ARG, "each family needs at least one element");
    for (int j = 0; j < n[k]; ++j) {
      const int e = fam[k][j];
      if (e < 0 || e >= (int)R.e.size()) return fail(ACCP_E_ARG, "family element index out of range");
      accp_elem c = R.e[e];
      if (!knob_field(c, quantity))
        return fail(ACCP_E_ARG, std::string("element ") + std::to_string(e) +
                                    (quantity == ACCP_TUNE ? " is neither a Quad nor a StrMPole"
                                                           : " is not a StrMPole"));
    }
  }
  return ACCP_OK;
}

// aline.cpp:1187-1242 / :1310-1355: +1e-5 on every member, measure, restore.
int resp_fd(accp_lattice *L, Records &R, int quantity, const int32_t *fam[2], const int n[2], double *resp) {
  double a0, b0;
  int rc = measure(L, quantity, &a0, &b0);
  for (int k = 0; k < 2 && !rc; ++k) {
    std::vector<double> old(n[k]);
    for (int j = 0; j < n[k]; ++j) {
      double *f = knob_field(R.e[fam[k][j]], quantity);
      old[j] = *f;
      *f = old[j] + DK_RESP;
    }
    double a = 0, b = 0;
    if (!(rc = R.store(L))) rc = measure(L, quantity, &a, &b);
    resp[k * 2 + 0] = (a - a0) / DK_RESP;
    resp[k * 2 + 1] = (b - b0) / DK_RESP;
    for (int j = 0; j < n[k]; ++j) *knob_field(R.e[fam[k][j]], quantity) = old[j];
    const int rc2 = R.store(L);
    if (!rc) rc = rc2;
  }
  return rc;
}

// Exact derivative of the getm66-based tunes.  With T = M11 + M22 (M33 + M44)
// a difference of ray coordinates over 2 delta, tune = acos(T/2)/2pi gives
// dtune/dk = -(dT/dk) / (4 pi sin(2 pi tune)); dT/dk is one backward pass of
// the turn map seeded with +-1/(2 delta) on the four rays that form T.
// Chromaticity differentiates the two tunes at dp -+ 1e-5 the same way.
int resp_adjoint(accp_lattice *L, const Records &R, int quantity, const int32_t *fam[2], const int n[2],
                 double *resp) {
  using accp::host::Knob;
  std::vector<Knob> knobs;
  for (int k = 0; k < 2; ++k) {
    const bool quad = R.e[fam[k][0]].type == ACCP_QUAD;
    for (int j = 1; j < n[k]; ++j)
      if ((R.e[fam[k][j]].type == ACCP_QUAD) != quad)
        return fail(ACCP_E_ARG, "a family mixes Quad and StrMPole elements");
    Knob kn;
    kn.field = quad ? ACCP_FIELD_K : ACCP_FIELD_POLYB;
    kn.index = quad ? 0 : (quantity == ACCP_TUNE ? 1 : 2);
    kn.members.assign(fam[k], fam[k] + n[k]);
    knobs.push_back(kn);
  }
  const double delta = ACCP_DELTA_DIFF;
  const int nset = quantity == ACCP_TUNE ? 1 : 2;  // ray sets: dp = 0, or dp = -1e-5 and +1e-5
  std::vector<double> r0(72 * nset);
  if (quantity == ACCP_TUNE) m66_rays(0.0, delta, r0.data());
  else {
    m66_rays(-DPP_CHROM, delta, r0.data());
    m66_rays(DPP_CHROM, delta, r0.data() + 72);
  }
  try {
    accp::host::TurnMapFunction F(L, knobs);
    const std::vector<double> out = F.forward(r0, 1, 1);
    double dtune[2][2][2];  // [ray set][plane][knob]
    for (int s = 0; s < nset; ++s) {
      double m[36], t[2];
      m66_from_rays(out.data() + 72 * s, delta, m);
      tunes_from_m66(m, &t[0], &t[1]);
      for (int p = 0; p < 2; ++p) {
        const int c = 2 * p;  // x or y
        std::vector<double> lam(out.size(), 0.0);
        double *l = lam.data() + 72 * s;
        l[c * 6 + c] = 1.0 / (2.0 * delta);
        l[36 + c * 6 + c] = -1.0 / (2.0 * delta);
        l[(c + 1) * 6 + c + 1] = 1.0 / (2.0 * delta);
        l[36 + (c + 1) * 6 + c + 1] = -1.0 / (2.0 * delta);
        const accp::host::Gradient g = F.backward(lam);
        const double sn = std::sin(2.0 * PI * t[p]);
        for (int k = 0; k < 2; ++k) dtune[s][p][k] = -g.dparam[k] / (4.0 * PI * sn);
      }
    }
    for (int k = 0; k < 2; ++k)
      for (int p = 0; p < 2; ++p)
        resp[k * 2 + p] = quantity == ACCP_TUNE ? dtune[0][p][k]
                                                : (dtune[1][p][k] - dtune[0][p][k]) / DPP_CHROM / 2.0;
  } catch (const std::exception &ex) {
    return fail(ACCP_E_LATTICE, ex.what());
  }
  return ACCP_OK;
}

// Solve A x = b (n <= 6) by Gaussian elimination with partial pivoting.
void solve_small(int n, double *A, double *b) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double sum = b[r];
    for (int k = r + 1; k < n; ++k) sum -= A[r * n + k] * b[k];
    b[r] = sum / A[r * n + r];
  }
}

// ALine::set_rad_onoff + set_cavity_onoff on the host records.
int set_flags(accp_lattice *L, Records &R, int rad, int cav) {
  for (auto &e : R.e) {
    if (e.type == ACCP_BENDMPOLE) e.rad = rad;
    if (e.type == ACCP_CAVITY) e.on = cav;
  }
  return R.store(L);
}

// Closed orbit rc = (I - M)^-1 r1 of the leading n x n block, r1 the
// one-turn image of (0,0,0,0,0,dp).
int closed_orbit(accp_lattice *L, int n, double dp, double *rc) {
  double m[36], r[6] = {0, 0, 0, 0, 0, dp}, A[36];
  int rc_ = accp_getm66(L, 0, -1, dp, ACCP_DELTA_DIFF, m);
  accp_track_opts o{};
  if (!rc_) rc_ = accp_track(L, r, 1, 1, 0, -1, &o);
  if (rc_) return rc_;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) A[i * n + j] = (i == j ? 1.0 : 0.0) - m[i * 6 + j];
    rc[i] = r[i];
  }
  solve_small(n, A, rc);
  return ACCP_OK;
}

// Least-squares slope of y on x, centred on the mean of x.
double lsq_slope(int n, const double *x, const double *y) {
  double sx = 0;
  for (int i = 0; i < n; ++i) sx += x[i];
  const double mx = sx / n;
  double stt = 0, sty = 0;
  for (int i = 0; i < n; ++i) {
    const double t = x[i] - mx;
    stt += t * t;
    sty += t * y[i];
  }
  return sty / stt;
}

}  // namespace

extern "C" int accp_getm66(accp_lattice *L, int st, int ed, double dp, double delta, double m66[36]) {
  if (!L || !m66) return fail(ACCP_E_ARG, "accp_getm66: null argument");
  if (!(delta > 0)) return fail(ACCP_E_ARG, "accp_getm66: delta must be positive");
  double r0[72];
  m66_rays(dp, delta, r0);
  accp_track_opts o{};
  int rc = accp_track(L, r0, 12, 1, st, ed, &o);
  if (rc) return rc;
  m66_from_rays(r0, delta, m66);
  return ACCP_OK;
}

extern "C" int accp_calctune(accp_lattice *L, double dp, double delta, double tune[2]) {
  if (!tune) return fail(ACCP_E_ARG, "accp_calctune: null argument");
  double m[36];
  int rc = accp_getm66(L, 0, -1, dp, delta, m);
  if (rc) return rc;
  tunes_from_m66(m, &tune[0], &tune[1]);
  return ACCP_OK;
}

extern "C" int accp_calcchrom(accp_lattice *L, double dp, double delta, double chrom[2]) {
  if (!chrom) return fail(ACCP_E_ARG, "accp_calcchrom: null argument");
  double m[36], tx0, ty0, tx, ty;
  int rc = accp_getm66(L, 0, -1, dp - DPP_CHROM, delta, m);
  if (rc) return rc;
  tunes_from_m66(m, &tx0, &ty0);
  if ((rc = accp_getm66(L, 0, -1, dp + DPP_CHROM, delta, m))) return rc;
  tunes_from_m66(m, &tx, &ty);
  chrom[0] = (tx - tx0) / DPP_CHROM / 2.0;
  chrom[1] = (ty - ty0) / DPP_CHROM / 2.0;
  return ACCP_OK;
}

extern "C" int accp_calcresp(accp_lattice *L, int quantity, int method, const int32_t *fam1, int n1,
                             const int32_t *fam2, int n2, double resp[4]) {
  if (!L || !resp) return fail(ACCP_E_ARG, "accp_calcresp: null argument");
  Records R;
  int rc = R.load(L);
  const int32_t *fam[2] = {fam1, fam2};
  const int n[2] = {n1, n2};
  if (!rc) rc = check_families(R, quantity, fam, n);
  if (rc) return rc;
  if (method == ACCP_RESP_FD) return resp_fd(L, R, quantity, fam, n, resp);
  if (method == ACCP_RESP_ADJOINT) return resp_adjoint(L, R, quantity, fam, n, resp);
  return fail(ACCP_E_ARG, "accp_calcresp: method must be ACCP_RESP_FD or ACCP_RESP_ADJOINT");
}

extern "C" int accp_correct(accp_lattice *L, int quantity, int method, const int32_t *fam1, int n1,
                            const int32_t *fam2, int n2, double target[2]) {
  if (!L || !target) return fail(ACCP_E_ARG, "accp_correct: null argument");
  // aline.cpp:1243-1309 / :1356-1418
  double a0, b0, resp[4];
  int rc = measure(L, quantity, &a0, &b0);
  if (!rc) rc = accp_calcresp(L, quantity, method, fam1, n1, fam2, n2, resp);
  if (rc) return rc;
  const double detr = resp[0] * resp[3] - resp[1] * resp[2];
  const double d1 = -(resp[3] * (a0 - target[0]) - resp[2] * (b0 - target[1])) / detr;
  const double d2 = -(-resp[1] * (a0 - target[0]) + resp[0] * (b0 - target[1])) / detr;
  Records R;
  if ((rc = R.load(L))) return rc;
  for (int j = 0; j < n1; ++j) *knob_field(R.e[fam1[j]], quantity) += d1;
  for (int j = 0; j < n2; ++j) *knob_field(R.e[fam2[j]], quantity) += d2;
  if ((rc = R.store(L))) return rc;
  return measure(L, quantity, &target[0], &target[1]);
}

extern "C" int accp_setrfphase(accp_lattice *L, double *phis) {
  if (!L || !phis) return fail(ACCP_E_ARG, "accp_setrfphase: null argument");
  Records R;
  int rc = R.load(L);
  if (rc) return rc;
  // aline.cpp:468-474: cavities off, classical radiation, one turn from the
  // origin.  A LongWake is off for the probe: the lone probe particle would
  // see its own field, the whole bunch charge's q k (0.31 MeV for 2 nC at
  // 5 kOhm, 10 GHz, Q = 1), which is not the synchronous loss.  The
  // reference's only wake, ATransWake, cannot kick a lone particle
  // (accel.cpp:1638), so its probe sees no wake either.
  std::vector<int32_t> wake_on;
  for (auto &e : R.e) {
    if (e.type == ACCP_CAVITY) e.on = 0;
    if (e.type == ACCP_BENDMPOLE) e.rad = ACCP_RAD_DAMP;
    if (e.type == ACCP_LONGWAKE) {
      wake_on.push_back(e.on);
      e.on = 0;
    }
  }
  if ((rc = R.store(L))) return rc;
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  accp_track_opts o{};
  if ((rc = accp_track(L, r, 1, 1, 0, -1, &o))) return rc;
  const double delta = -r[5];
  size_t iw = 0;
  for (auto &e : R.e) {
    if (e.type == ACCP_BENDMPOLE) e.rad = ACCP_RAD_OFF;
    if (e.type == ACCP_LONGWAKE) e.on = wake_on[iw++];
  }
  // aline.cpp:477-499
  double tot_volt = 0, energy = 0;
  int ncav = 0;
  for (const auto &e : R.e)
    if (e.type == ACCP_CAVITY) {
      tot_volt += e.volt;
      energy = e.energy;
      ++ncav;
    }
  *phis = 0.0;
  if (ncav) {
    *phis = std::asin(delta * energy / tot_volt);
    for (auto &e : R.e)
      if (e.type == ACCP_CAVITY) e.phis = *phis;
  }
  return R.store(L);
}

extern "C" int accp_trackda(accp_lattice *L, double xmax, double xmin, double ymax, int nturn, int nray,
                            int nstep, int negweight, double *rx, double *ry, double *area,
                            int32_t *loss_turn) {
  if (!L || !rx || !ry || !area) return fail(ACCP_E_ARG, "accp_trackda: null argument");
  if (!(xmax > 0) || !(xmin > 0) || !(ymax > 0))  // the reference asserts these (aline.cpp:610-612)
    return fail(ACCP_E_ARG, "accp_trackda: xmax, xmin and ymax must be positive");
  if (nturn < 0) return fail(ACCP_E_ARG, "accp_trackda: nturn must not be negative");
  // aline.cpp:615-629
  if (nray < 3) nray = 3;
  if (nray > 51) nray = 51;
  if (nstep < 2) nstep = 2;
  const double deltax = xmax - xmin;
  if (deltax < 0) return fail(ACCP_E_ARG, "xmax must be greater than xmin");
  const int np = nray * nstep;
  std::vector<double> xl(np), yl(np), r((size_t)np * 6, 0.0);
  for (int i = 0; i < nray; ++i)
    for (int j = 0; j < nstep; ++j) {
      const int q = i * nstep + j;
      xl[q] = (xmin + deltax * j / (nstep - 1)) * std::cos(i * PI / (nray - 1));
      yl[q] = (xmin + deltax * j / (nstep - 1)) * std::sin(i * PI / (nray - 1)) * ymax / xmax;
      r[q * 6 + 0] = xl[q];
      r[q * 6 + 2] = yl[q];
      r[q * 6 + 4] = 1.0e-5;
    }
  accp_track_opts o{};
  o.loss_turn = loss_turn;
  int rc = accp_track(L, r.data(), np, nturn, 0, -1, &o);
  if (rc) return rc;
  // aline.cpp:676-686
  for (int i = 0; i < nray; ++i) {
    rx[i] = xl[i * nstep + nstep - 1];
    ry[i] = yl[i * nstep + nstep - 1];
    for (int j = 0; j < nstep; ++j)
      if (std::isnan(r[(i * nstep + j) * 6])) {
        rx[i] = xl[i * nstep + j];
        ry[i] = yl[i * nstep + j];
        break;
      }
  }
  // aline.cpp:688-707
  double s = 0;
  for (int i = 1; i < nray; ++i) {
    const double rd1 = std::sqrt(rx[i - 1] * rx[i - 1] + ry[i - 1] * ry[i - 1]);
    const double rd2 = std::sqrt(rx[i] * rx[i] + ry[i] * ry[i]);
    const double dphi = std::atan2(ry[i], rx[i]) - std::atan2(ry[i - 1], rx[i - 1]);
    if (rx[i] <= 0) s += .5 * rd1 * rd2 * std::sin(dphi);
    else if (negweight) s += .25 * rd1 * rd2 * std::sin(dphi);
    else s += .5 * rd1 * rd2 * std::sin(dphi);
  }
  *area = s;
  return ACCP_OK;
}

extern "C" int accp_getco4(accp_lattice *L, double dp, double rc[4]) {
  if (!L || !rc) return fail(ACCP_E_ARG, "accp_getco4: null argument");
  Records R;
  int rc_ = R.load(L);
  if (!rc_) rc_ = set_flags(L, R, ACCP_RAD_OFF, 0);  // aline.cpp:841-842
  if (rc_) return rc_;
  return closed_orbit(L, 4, dp, rc);
}

extern "C" int accp_getco6(accp_lattice *L, double rc[6]) {
  if (!L || !rc) return fail(ACCP_E_ARG, "accp_getco6: null argument");
  Records R;
  int rc_ = R.load(L);
  if (!rc_) rc_ = set_flags(L, R, ACCP_RAD_DAMP, 1);  // aline.cpp:874-875
  if (rc_) return rc_;
  return closed_orbit(L, 6, 0.0, rc);
}

extern "C" int accp_calctwiss(accp_lattice *L, const int32_t *idx, int n, double dp, double delta,
                              double *tw) {
  if (!L || !tw || n < 0 || (n > 0 && !idx)) return fail(ACCP_E_ARG, "accp_calctwiss: bad arguments");
  Records R;
  int rc = R.load(L);
  if (rc) return rc;
  const int ne = (int)R.e.size();
  for (int i = 0; i < n; ++i)
    if (idx[i] < 0 || idx[i] > ne || (i && idx[i] < idx[i - 1]))
      return fail(ACCP_E_ARG, "accp_calctwiss: indices must be increasing and within the line");
  // aline.cpp:1097-1106: periodic functions from the one-turn matrix
  double m[36];
  if ((rc = accp_getm66(L, 0, -1, dp, delta, m))) return rc;
  double t0[7];
  const double psix = std::acos((m[0] + m[7]) * 0.5), psiy = std::acos((m[14] + m[21]) * 0.5);
  t0[0] = 0;
  t0[1] = psix;
  t0[2] = psiy;
  t0[3] = m[1] / std::sin(psix);
  t0[4] = (m[0] - m[7]) / std::sin(psix) / 2.0;
  t0[5] = m[15] / std::sin(psiy);
  t0[6] = (m[14] - m[21]) / std::sin(psiy) / 2.0;
  if (n == 0) {
    for (int k = 0; k < 7; ++k) tw[k] = t0[k];
    return ACCP_OK;
  }
  // aline.cpp:1057-1093: carried from point to point
  t0[1] = t0[2] = 0;
  auto length = [&](int st, int ed) {
    double l = 0;
    for (int k = st; k < ed; ++k) l += R.e[k].len;
    return l;
  };
  int id = idx[0];
  double T[36];
  if ((rc = accp_getm66(L, 0, id, dp, delta, T))) return rc;
  double len01 = length(0, id);
  for (int i = 0; i < n; ++i) {
    const double bx = t0[3], ax = t0[4], by = t0[5], ay = t0[6];
    double *t1 = tw + 7 * i;
    auto sq = [](double v) { return v * v; };
    t1[0] = t0[0] + len01;
    t1[3] = 1.0 / bx * (sq(T[0] * bx - T[1] * ax) + sq(T[1]));
    t1[5] = 1.0 / by * (sq(T[14] * by - T[15] * ay) + sq(T[15]));
    t1[4] = -1.0 / bx * ((T[0] * bx - T[1] * ax) * (T[6] * bx - T[7] * ax) + T[1] * T[7]);
    t1[6] = -1.0 / by * ((T[14] * by - T[15] * ay) * (T[20] * by - T[21] * ay) + T[15] * T[21]);
    t1[1] = t0[1] + std::atan2(T[1], T[0] * bx - T[1] * ax);
    t1[2] = t0[2] + std::atan2(T[15], T[14] * by - T[15] * ay);
    for (int k = 0; k < 7; ++k) t0[k] = t1[k];
    if (i < n - 1) {
      const int id2 = idx[i + 1];
      if ((rc = accp_getm66(L, id, id2, dp, delta, T))) return rc;
      len01 = length(id, id2);
      id = id2;
    }
  }
  return ACCP_OK;
}

extern "C" int accp_calcchrom2(accp_lattice *L, double c[4]) {
  if (!L || !c) return fail(ACCP_E_ARG, "accp_calcchrom2: null argument");
  // aline.cpp:1133-1157
  double cx[11], cy[11], dp[11], v[2];
  for (int i = 0; i < 11; ++i) {
    dp[i] = 1.0e-3 * (i - 6.0);
    const int rc = accp_calcchrom(L, dp[i], ACCP_DELTA_DIFF, v);
    if (rc) return rc;
    cx[i] = v[0];
    cy[i] = v[1];
  }
  c[2] = lsq_slope(11, dp, cx) / 2.0;
  c[3] = lsq_slope(11, dp, cy) / 2.0;
  const int rc = accp_calcchrom(L, 0.0, ACCP_DELTA_DIFF, v);
  c[0] = v[0];
  c[1] = v[1];
  return rc;
}

extern "C" int accp_trackma(accp_lattice *L, double dppmax, double dppmin, int nturn, int nstep,
                            const int32_t *locs, int nloc, double *dpp_plus, double *dpp_minus, double *avg) {
  if (!L || !locs || nloc <= 0 || !dpp_plus || !dpp_minus || !avg)
    return fail(ACCP_E_ARG, "accp_trackma: bad arguments");
  if (!(dppmax > 0) || !(dppmin > 0) || nstep < 1 || nturn < 0)  // aline.cpp:783-784
    return fail(ACCP_E_ARG, "accp_trackma: dppmax and dppmin must be positive, nstep at least 1");
  Records R;
  int rc = R.load(L);
  if (rc) return rc;
  const int ne = (int)R.e.size();
  for (int k = 0; k < nloc; ++k)
    if (locs[k] < 0 || locs[k] >= ne) return fail(ACCP_E_ARG, "accp_trackma: location out of range");
  // One lattice per location, the ring started there (resetfirstelement,
  // aline.cpp:557-606), all tracked in one batched launch.
  std::vector<accp_lattice *> lats(nloc, nullptr);
  accp_batch *b = nullptr;
  double *d_r = nullptr;
  const int np = 2 * (nstep + 1);
  const size_t bytes = sizeof(double) * 6 * (size_t)np * nloc;
  std::vector<double> r((size_t)6 * np * nloc, 0.0);
  std::vector<int64_t> offs(nloc + 1);
  const double stepsize = (dppmax - dppmin) / nstep;
  for (int k = 0; k < nloc && !rc; ++k) {
    std::vector<accp_elem> rot(ne);
    for (int i = 0; i < ne; ++i) rot[i] = R.e[(locs[k] + i) % ne];
    rc = accp_lattice_build(rot.data(), ne, &lats[k]);
    double *q = r.data() + (size_t)6 * np * k;
    for (int j = 0; j < nstep + 1; ++j) {  // aline.cpp:797-801
      q[j * 12 + 5] = dppmin + stepsize * j;
      q[j * 12 + 6 + 5] = -(dppmin + stepsize * j);
    }
    offs[k] = (int64_t)np * k;
  }
  offs[nloc] = (int64_t)np * nloc;
  if (!rc) rc = accp_batch_create(lats.data(), nloc, &b);
  if (!rc && hipMalloc(&d_r, bytes) != hipSuccess) rc = fail(ACCP_E_NOMEM, "accp_trackma: hipMalloc failed");
  if (!rc && hipMemcpy(d_r, r.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ACCP_E_HIP, "accp_trackma: upload failed");
  if (!rc) rc = accp_track_batched(b, d_r, offs.data(), nturn, nullptr, nullptr, nullptr);
  if (!rc && (hipDeviceSynchronize() != hipSuccess ||
              hipMemcpy(r.data(), d_r, bytes, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(ACCP_E_HIP, "accp_trackma: tracking or download failed");
  if (d_r) (void)hipFree(d_r);
  accp_batch_free(b);
  for (accp_lattice *l : lats) accp_lattice_free(l);
  if (rc) return rc;
  // aline.cpp:818-831: the largest offset that survives, on each side
  double sum = 0;
  for (int k = 0; k < nloc; ++k) {
    const double *q = r.data() + (size_t)6 * np * k;
    dpp_plus[k] = dppmin;
    dpp_minus[k] = -dppmin;
    for (int j = 1; j < nstep + 1; ++j) {
      if (!std::isnan(q[j * 12])) dpp_plus[k] = dppmin + stepsize * j;
      if (!std::isnan(q[j * 12 + 6])) dpp_minus[k] = -(dppmin + stepsize * j);
    }
    sum += std::fmin(dpp_plus[k], -dpp_minus[k]);
  }
  *avg = sum / nloc;
  return ACCP_OK;
}
