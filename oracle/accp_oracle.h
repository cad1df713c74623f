/*
 * accp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the Accel+ tracking formulas, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (the accp C ABI, its HIP kernels, its host driver)
 * links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the Accel+ tree).  Built with -O2 -ffp-contract=off and no -march, the
 * same arithmetic the reference gets from g++ -O2 on x86-64 (no FMA).
 */
#ifndef ACCP_ORACLE_H
#define ACCP_ORACLE_H

#include <stdint.h>
#include "../include/accp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- integrator primitives on one r[6] (src/lib/accel/atpass.cpp) ------ */
void orc_drift_kzero(double *r, double L);                            /* :6-13    */
void orc_quad6(double *r, double L, double K);                        /* :15-72   */
void orc_bend6(double *r, double L, double angle, double grd, double by_err); /* :76-187 */
void orc_edge(double *r, double inv_rho, double edge_angle);          /* :190-196 */
void orc_bndthinkick(double *r, const double *B, double L, double irho, int max_order); /* :199-235 */
double orc_b2perp(double bx, double by, double irho, double x, double xpr,
                  double y, double ypr);                              /* :237-253 */
void orc_bndthinkickrad(double *r, const double *B, double L, double irho,
                        double E0, int max_order, int flag_rad, double gauss); /* :255-339 */
void orc_hxdrift6(double *r, double L, double h);                     /* :343-355 */
void orc_strthinkick(double *r, const double *A, const double *B, double L,
                     int max_order);                                  /* :358-414 */
void orc_fastdrift(double *r, double norm_l);                         /* :416-428 */

/* ---- counter-based Gaussian stream (quantum excitation) ----------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Standard normal for one radiation kick.  Counter = (particle id, turn,
 * element index, slice / 2), key = seed; Philox4x32-10 -> two 53-bit
 * uniforms u1, u2 in (0,1) -> Box-Muller, sqrt(-2 ln u1) times cos(2 pi u2)
 * for an even slice and sin(2 pi u2) for an odd one (two slices share one
 * Philox block). */
double orc_gauss(uint64_t seed, uint64_t pid, uint64_t turn, uint32_t eidx,
                 uint32_t slice);

/* ---- one element over a beam (src/lib/accel/accel.cpp) ------------------ */
typedef struct orc_ctx {
  uint64_t seed;  /* RNG key                                  */
  uint64_t turn;  /* absolute turn number                     */
  uint64_t pid0;  /* particle id of r[0]                      */
  int32_t eidx;   /* element index in the line                */
} orc_ctx;

void orc_elem_pass(const accp_elem *e, double *r, int n, const orc_ctx *ctx);

/* ---- line (src/lib/accel/aline.cpp:313-363) ----------------------------- */
/* Elements [st, ed) on all particles, nturn times; ed < 0 = whole line. */
void orc_line_pass(const accp_elem *elems, int nelem, double *r, int n, int st,
                   int ed, int nturn, uint64_t seed, uint64_t turn0, uint64_t pid0);
/* As orc_line_pass, with OpenMP over contiguous particle blocks the way
 * ALine::pass(r,N,st,ed) partitions them (aline.cpp:313-329).  Elements
 * that need the whole beam (LongWake) are rejected (returns -1). */
int orc_line_pass_omp(const accp_elem *elems, int nelem, double *r, int n,
                      int nturn, int nthreads, uint64_t seed, uint64_t turn0, uint64_t pid0);

/* ---- optics by tracking (src/lib/accel/aline.cpp) ----------------------- */
void orc_getm66(const accp_elem *elems, int nelem, int st, int ed, double dpp,
                double delta, double m66[36]);                        /* :900-927  */
void orc_calctune(const accp_elem *elems, int nelem, double dpp, double delta,
                  double *tunex, double *tuney);                      /* :1032-1043 */
void orc_calcchrom(const accp_elem *elems, int nelem, double dp, double delta,
                   double *chromx, double *chromy);                   /* :1108-1126 */
/* Knob kinds of calctuneresp / calcchromresp. */
#define ORC_KNOB_TUNE 1  /* Quad K or StrMPole PolyB[1] (+1e-5)  */
#define ORC_KNOB_CHROM 2 /* StrMPole PolyB[2] (+1e-5)            */
void orc_calcresp(accp_elem *elems, int nelem, int kind, const int *fam1, int n1,
                  const int *fam2, int n2, double resp[4]);           /* :1187-1242, :1310-1355 */
/* One correction step; target in/out as in ALine::correcttune/correctchrom
 * (:1243-1309, :1356-1418).  Modifies elems in place. */
void orc_correct(accp_elem *elems, int nelem, int kind, const int *fam1, int n1,
                 const int *fam2, int n2, double *tx, double *ty);

/* ---- studies ------------------------------------------------------------ */
/* ALine::setrfphase (aline.cpp:465-501): returns phis, writes it into every
 * cavity; flags are restored the way the reference leaves them
 * (radiation off, cavities off). */
double orc_setrfphase(accp_elem *elems, int nelem);
/* ALine::trackDA without the loss-turn printout (aline.cpp:607-715). */
double orc_trackda(const accp_elem *elems, int nelem, double xmax, double xmin,
                   double ymax, int nturn, int nray, int nstep, double *rx,
                   double *ry, int negweight);
/* ALine::getco4 (aline.cpp:838-870): 4D closed orbit at dp with radiation
 * and cavities off (the flags are left so, as the reference leaves them):
 * rc = (I - M44)^-1 r1 with r1 the one-turn image of (0,0,0,0,0,dp).  The
 * reference inverts by SVD (matvec.cpp:269-284); this solves the same
 * system by Gaussian elimination with partial pivoting. */
void orc_getco4(accp_elem *elems, int nelem, double dp, double rc[4]);
/* ALine::getco6 (aline.cpp:871-889): 6D, radiation (classical) and
 * cavities on, rc = (I - M66)^-1 r1 from the origin. */
void orc_getco6(accp_elem *elems, int nelem, double rc[6]);
/* ALine::calctwiss at the ring start (aline.cpp:1095-1107) and propagated
 * to elements idx[0..n) (aline.cpp:1044-1094); per point
 * {s, psix, psiy, betx, alfx, bety, alfy}. */
void orc_calctwiss0(const accp_elem *elems, int nelem, double dp, double tw[7]);
void orc_calctwiss(const accp_elem *elems, int nelem, const int *idx, int n, double dp,
                   double *tw);
/* ALine::calcchrom2 (aline.cpp:1127-1158): chromaticities at
 * dpp = 1e-3 (i - 6), i = 0..10, a least-squares line through them
 * (slope / 2 = second-order chromaticity); c = {chromx, chromy, chromx2,
 * chromy2}.  The line is the textbook centred least-squares slope. */
void orc_calcchrom2(const accp_elem *elems, int nelem, double c[4]);
/* ALine::trackMA (aline.cpp:781-836): for each location, the ring started
 * there, 2 (nstep+1) particles at +-(dppmin + j step), nturn turns; the
 * largest surviving offset on each side; returns the average of
 * min(dpp+, -dpp-). */
double orc_trackma(const accp_elem *elems, int nelem, double dppmax, double dppmin, int nturn,
                   int nstep, const int *locs, int nloc, double *dpp_plus, double *dpp_minus);

/* ---- tune diagnostics ------------------------------------------------- */
/* naff (utility.cpp:48-96): the frequency (tune, 0..0.5) of the strongest
 * line of data[0..n) between xlow and xhi; n a power of two. */
double orc_naff(int n, const double *data, double xlow, double xhi);
/* fprojection (utility.cpp:21-41). */
double orc_fprojection(double nv, int n, const double *x);
/* ALine::tracktune (aline.cpp:1159-1186). */
void orc_tracktune(const accp_elem *elems, int nelem, unsigned nturn, double x0, double y0, double xlow,
                   double xhi, double ylow, double yhi, double tune[2]);
/* ALine::trackDiffusion (aline.cpp:1421-1484); returns the sum of dif. */
double orc_trackdiffusion(const accp_elem *elems, int nelem, int nturn, int np, const double *x0,
                          const double *y0, double *nux, double *nuy, double *dif, double xlow, double xhi,
                          double ylow, double yhi);
/* ALine::trackdnudexy (aline.cpp:716-779). */
void orc_trackdnudexy(const accp_elem *elems, int nelem, unsigned nturn, double xm, double ym, double d[4]);

/* calc_beam_stat (utility.cpp:322-373) into an ACCP_STAT_LEN record. */
int orc_beam_stat(const double *r, int n, double turn, double *stat);

/* ---- longitudinal resonator wake (new element; closed form) ------------- */
/* W(tau) of a resonator: 2*alpha*Rs*exp(-alpha*tau)*(cos(wb*tau) -
 * alpha/wb*sin(wb*tau)) for tau > 0 and alpha*Rs at tau = 0. */
double orc_wake_w(double rs, double q, double fr, double tau);
/* Brute-force O(N^2) energy kick of every live particle from all others
 * (no binning): used to check the binned element. */
void orc_wake_pairwise(const accp_elem *e, const double *r, int n, double *ddelta);

#ifdef __cplusplus
}
#endif
#endif
