// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A small extern "C" shim over the compiled Accel+ reference library (its
// own sources, unmodified, compiled by oracle/Makefile into
// oracle/_ref/libaccel_ref.so).  Tests load it with ctypes to write golden
// fixtures and to pin the oracle; bench.py times ref_pass as the CPU
// baseline.  Nothing in the product links it.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include <omp.h>

#include "accel.h"
#include "matvec.h"
#include "utility.h"

namespace {
// Common random numbers for quantum excitation.  The library is linked with
// -Wl,--wrap=rand, so the reference's grandn (utility.cpp:194-217) draws its
// uniforms from __wrap_rand below.  Outside ref_pass_crn that forwards to
// the C library; inside it, while one particle passes one element that draws
// normals (BendMPole with quantum excitation, SyncRad, CrabCav, RFMult),
// call c returns the oracle's uniform u1 (c even) or u2 (c odd) of pair c/2
// of that (particle, turn, element): Philox4x32-10 (Salmon et al., SC'11)
// keyed by the seed over the counter {pid, turn, element, pair}, the words
// taken as the oracle's 53-bit uniform, rounded to rand()'s grid
// k/RAND_MAX.  The reference's Box-Muller then yields the oracle's normals to
// ~1e-9: the j-th normal a particle draws in an element is the cosine (j
// even) or sine (j odd) branch of pair j/2.  grandn keeps the second normal
// of a pair in static state and flips a static toggle per call
// (utility.cpp:194-217), so after each particle the harness brings the
// toggle back to "draw a new pair" (sync_grandn).
struct Crn {
  bool on = false;
  uint32_t key[2] = {0, 0};
  uint32_t pid = 0, turn = 0, eidx = 0, calls = 0;
} g_crn;

void crn_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Leave grandn's toggle at "draw a new pair": one call draws (two rand()
// calls) when it was there, and then a second call flips it back; a call
// that draws nothing consumed the pending second normal and already did.
void sync_grandn() {
  const uint32_t before = g_crn.calls;
  grandn();
  if (g_crn.calls != before) grandn();
}

bool draws_normals(AccelElement *el) {
  const std::string &t = el->gettype();
  return t == "BendMPole" || t == "SyncRad" || t == "CrabCav" || t == "RFMult";
}

bool g_map_ready = false;
void ensure_map() {
  if (!g_map_ready) {
    AccelElement::initializemap();
    g_map_ready = true;
  }
}
}  // namespace

extern "C" int __real_rand(void);
extern "C" int __wrap_rand(void) {
  if (!g_crn.on) return __real_rand();
  const uint32_t c = g_crn.calls++;
  uint32_t w[4] = {g_crn.pid, g_crn.turn, g_crn.eidx, c >> 1};
  crn_philox(w, g_crn.key[0], g_crn.key[1]);
  const uint32_t a = (c & 1) ? w[2] : w[0], b = (c & 1) ? w[3] : w[1];
  const uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  const double u = ((double)k + 0.5) * (1.0 / 9007199254740992.0);
  long v = std::lround(u * RAND_MAX);
  // u1 = 0 would make grandn draw again (utility.cpp:206-210)
  if (v < 1) v = 1;
  if (v > RAND_MAX) v = RAND_MAX;
  return (int)v;
}

extern "C" {

// ALine::deserialize on a lattice file (aline.cpp:202-219).
void *ref_load(const char *path) {
  ensure_map();
  std::ifstream fin(path);
  if (!fin.good()) return nullptr;
  ALine *line = ALine::deserialize(fin);
  return line;
}

void ref_free(void *h) { delete static_cast<ALine *>(h); }

int ref_nelem(void *h) { return (int)static_cast<ALine *>(h)->getnumber(); }

void ref_set_rad(void *h, int flag) { static_cast<ALine *>(h)->set_rad_onoff((unsigned)flag); }

void ref_set_cavity(void *h, int on) { static_cast<ALine *>(h)->set_cavity_onoff(on != 0); }

void ref_set_wake(void *h, int on) { static_cast<ALine *>(h)->set_wake_onoff(on != 0); }

void ref_set_crab(void *h, int on) { static_cast<ALine *>(h)->set_crab_onoff(on != 0); }

void ref_set_threads(int n) { omp_set_num_threads(n < 1 ? 1 : n); }

// ALine::pass(r, N, nturn) (aline.cpp:359-363).
void ref_pass(void *h, double *r, int n, int nturn) { static_cast<ALine *>(h)->pass(r, n, nturn); }

// ALine::pass(r, N, st, ed) (aline.cpp:313-329).
void ref_pass_range(void *h, double *r, int n, int st, int ed) {
  static_cast<ALine *>(h)->pass(r, n, (unsigned)st, (unsigned)ed);
}

// One element's pass on its own (AccelElement::pass, accel.h:70).
void ref_elem_pass(void *h, int idx, double *r, int n) {
  static_cast<ALine *>(h)->getelement((unsigned)idx)->pass(r, n);
}

double ref_setrfphase(void *h) { return static_cast<ALine *>(h)->setrfphase(); }

void ref_getm66(void *h, double dpp, double delta, double *m66) {
  Mat m(6, 6);
  static_cast<ALine *>(h)->getm66(m, dpp, delta);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) m66[i * 6 + j] = m[i + 1][j + 1];
}

void ref_calctune(void *h, double *tx, double *ty) { static_cast<ALine *>(h)->calctune(*tx, *ty); }

void ref_calcchrom(void *h, double *cx, double *cy) { static_cast<ALine *>(h)->calcchrom(*cx, *cy); }

// kind 1: correcttune, kind 2: correctchrom; families looked up by type and
// name the way proc_lattice does (process.cpp:624-639, :737-751).
int ref_correct(void *h, int kind, const char *type, const char *fam1, const char *fam2,
                double *tx, double *ty) {
  ALine *line = static_cast<ALine *>(h);
  unsigned n = line->getnumber();
  unsigned *l1 = new unsigned[n], *l2 = new unsigned[n];
  unsigned n1 = line->findelements(l1, std::string(type), std::string(fam1));
  unsigned n2 = line->findelements(l2, std::string(type), std::string(fam2));
  if (n1 < 1 || n2 < 1) {
    delete[] l1;
    delete[] l2;
    return -1;
  }
  if (kind == 1) line->correcttune(*tx, *ty, n1, n2, l1, l2);
  else line->correctchrom(*tx, *ty, n1, n2, l1, l2);
  delete[] l1;
  delete[] l2;
  return 0;
}

// Tune / chromaticity response matrices (aline.cpp:1187-1242, :1310-1355).
int ref_calcresp(void *h, int kind, const char *type, const char *fam1, const char *fam2,
                 double *resp) {
  ALine *line = static_cast<ALine *>(h);
  unsigned n = line->getnumber();
  unsigned *l1 = new unsigned[n], *l2 = new unsigned[n];
  unsigned nf[2];
  nf[0] = line->findelements(l1, std::string(type), std::string(fam1));
  nf[1] = line->findelements(l2, std::string(type), std::string(fam2));
  unsigned *lf[2] = {l1, l2};
  double r[2][2];
  if (kind == 1) line->calctuneresp(r, 2, nf, lf);
  else line->calcchromresp(r, 2, nf, lf);
  resp[0] = r[0][0]; resp[1] = r[0][1]; resp[2] = r[1][0]; resp[3] = r[1][1];
  delete[] l1;
  delete[] l2;
  return 0;
}

double ref_trackda(void *h, double xmax, double xmin, double ymax, int nturn, int nray, int nstep,
                   double *rx, double *ry, int negweight) {
  return static_cast<ALine *>(h)->trackDA(xmax, xmin, ymax, nturn, (unsigned)nray, nstep, rx, ry,
                                          negweight != 0);
}

int ref_beam_stat(double *r, int n, double *ar, double *sig, double *sigrr) {
  return calc_beam_stat(r, n, ar, sig, sigrr);
}

// Element field setters (AccelElement::setfield and the indexed multipole
// variants, accel.cpp:210-219, :773-785).
void ref_setfield(void *h, int idx, const char *name, double v) {
  static_cast<ALine *>(h)->getelement((unsigned)idx)->setfield(&v, std::string(name));
}
void ref_setfield_int(void *h, int idx, const char *name, int v) {
  static_cast<ALine *>(h)->getelement((unsigned)idx)->setfield(&v, std::string(name));
}
void ref_set_mpole(void *h, int idx, const char *name, int index, double v) {
  AccelElement *e = static_cast<ALine *>(h)->getelement((unsigned)idx);
  if (e->gettype() == "StrMPole") static_cast<AStrMPole *>(e)->setfield(&v, std::string(name), index);
  else if (e->gettype() == "BendMPole") static_cast<ABendMPole *>(e)->setfield(&v, std::string(name), index);
}
double ref_getfield(void *h, int idx, const char *name) {
  double v = 0;
  static_cast<ALine *>(h)->getelement((unsigned)idx)->getfield(&v, std::string(name));
  return v;
}
double ref_get_mpole(void *h, int idx, const char *name, int index) {
  double v = 0;
  AccelElement *e = static_cast<ALine *>(h)->getelement((unsigned)idx);
  if (e->gettype() == "StrMPole") static_cast<AStrMPole *>(e)->getfield(&v, std::string(name), index);
  else if (e->gettype() == "BendMPole") static_cast<ABendMPole *>(e)->getfield(&v, std::string(name), index);
  return v;
}

// ALine::getco4 / getco6 (aline.cpp:838-889): closed orbit; both set the
// radiation and cavity flags as the reference does.
void ref_getco4(void *h, double dp, double *rc) {
  Vec v(4);
  static_cast<ALine *>(h)->getco4(v, dp);
  for (int i = 0; i < 4; i++) rc[i] = v[i + 1];
}
void ref_getco6(void *h, double *rc) {
  Vec v(6);
  static_cast<ALine *>(h)->getco6(v);
  for (int i = 0; i < 6; i++) rc[i] = v[i + 1];
}

// ALine::calctwiss at the ring start (aline.cpp:1095-1107) and at element
// indices (aline.cpp:1044-1094): per point {s, psix, psiy, betx, alfx, bety, alfy}.
void ref_calctwiss0(void *h, double dp, double *tw) {
  Twiss t;
  static_cast<ALine *>(h)->calctwiss(t, dp);
  const double v[7] = {t.s, t.psix, t.psiy, t.betx, t.alfx, t.bety, t.alfy};
  for (int i = 0; i < 7; i++) tw[i] = v[i];
}
void ref_calctwiss(void *h, const int *idx, int n, double dp, double *tw) {
  Twiss *t = new Twiss[n];
  unsigned *u = new unsigned[n];
  for (int i = 0; i < n; i++) u[i] = (unsigned)idx[i];
  static_cast<ALine *>(h)->calctwiss(t, u, n, dp);
  for (int i = 0; i < n; i++) {
    const double v[7] = {t[i].s, t[i].psix, t[i].psiy, t[i].betx, t[i].alfx, t[i].bety, t[i].alfy};
    for (int k = 0; k < 7; k++) tw[7 * i + k] = v[k];
  }
  delete[] t;
  delete[] u;
}

// ALine::calcchrom2 (aline.cpp:1127-1158)
void ref_calcchrom2(void *h, double *c) {
  static_cast<ALine *>(h)->calcchrom2(c[0], c[1], c[2], c[3]);
}

// ALine::trackMA (aline.cpp:781-836)
double ref_trackma(void *h, double dppmax, double dppmin, int nturn, int nstep, const int *locs, int n,
                   double *dpp_plus, double *dpp_minus) {
  unsigned *u = new unsigned[n];
  for (int i = 0; i < n; i++) u[i] = (unsigned)locs[i];
  const double avg = static_cast<ALine *>(h)->trackMA(dppmax, dppmin, nturn, (unsigned)nstep, dpp_plus,
                                                      dpp_minus, (unsigned)n, u);
  delete[] u;
  return avg;
}

// naff (utility.cpp:48-96) on a copy of the signal
double ref_naff(int n, const double *data, double xlow, double xhi) {
  double *d = new double[n];
  std::memcpy(d, data, sizeof(double) * n);
  const double v = naff(n, d, xlow, xhi);
  delete[] d;
  return v;
}

// ALine::tracktune (aline.cpp:1159-1186)
void ref_tracktune(void *h, unsigned nturn, double x0, double y0, double xlow, double xhi, double ylow,
                   double yhi, double *tune) {
  static_cast<ALine *>(h)->tracktune(tune[0], tune[1], nturn, x0, y0, xlow, xhi, ylow, yhi);
}

// ALine::trackDiffusion (aline.cpp:1421-1484)
double ref_trackdiffusion(void *h, int nturn, int np, double *x0, double *y0, double *nux, double *nuy,
                          double *dif, double xlow, double xhi, double ylow, double yhi) {
  return static_cast<ALine *>(h)->trackDiffusion(nturn, (unsigned)np, x0, y0, nux, nuy, dif, xlow, xhi, ylow,
                                                 yhi);
}

// ALine::trackdnudexy (aline.cpp:716-779)
void ref_trackdnudexy(void *h, unsigned nturn, double xm, double ym, double *d) {
  static_cast<ALine *>(h)->trackdnudexy(d, nturn, xm, ym);
}
// ALine::pass(r, N, nturn) (aline.cpp:359-363) with the oracle's normals in
// every noise draw (common random numbers, see __wrap_rand): elements in line
// order over the whole beam, except that each element that draws normals
// passes one particle at a time, so that its draws take the counter stream of
// (pid0 + i, turn0 + t, element).  Particles are independent (no collective
// element may be in the line), so the order changes nothing else.  Returns 0.
int ref_pass_crn(void *h, double *r, int n, int nturn, uint64_t seed, uint64_t turn0, uint64_t pid0) {
  ALine *line = static_cast<ALine *>(h);
  const unsigned ne = line->getnumber();
  g_crn.key[0] = (uint32_t)seed;
  g_crn.key[1] = (uint32_t)(seed >> 32);
  g_crn.on = true;
  sync_grandn();
  g_crn.on = false;
  for (int t = 0; t < nturn; ++t) {
    for (unsigned e = 0; e < ne; ++e) {
      AccelElement *el = line->getelement(e);
      if (!draws_normals(el)) {
        el->pass(r, n);
        continue;
      }
      for (int i = 0; i < n; ++i) {
        g_crn.pid = (uint32_t)(pid0 + (uint64_t)i);
        g_crn.turn = (uint32_t)(turn0 + (uint64_t)t);
        g_crn.eidx = e;
        g_crn.calls = 0;
        g_crn.on = true;
        el->pass(r + 6 * i, 1);
        sync_grandn();
        g_crn.on = false;
      }
    }
  }
  return 0;
}
}
