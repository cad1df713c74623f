/*
 * accp.h -- C ABI of the MI355X (gfx950) tracking backend for Accel+.
 *
 * Plain C, no C++ types, no exceptions across the boundary.  Every entry
 * point returns 0 on success and a negative ACCP_E* code on failure; the
 * message of the last failure on the calling thread is accp_last_error().
 *
 * Particle layout is the reference's own: an AoS double[6*N] buffer of
 * {x, px, y, py, z, delta} per particle, updated in place; a particle is lost
 * when x is NaN (then z holds the aperture loss code), exactly as
 * AccelElement::pass(double*, int) treats it (src/lib/accel/accel.h:70,
 * e.g. src/lib/accel/accel.cpp:166-167, :350-354).
 *
 * Reference interfaces replaced (file:line, paths relative to the Accel+ tree):
 *   accp_lattice_build / accp_lattice_update
 *        <- ALine::deserialize + AccelElement::create (src/lib/accel/aline.cpp:202-219,
 *           src/lib/accel/accel.cpp:67-133) and the flag/field setters
 *           ALine::set_rad_onoff / set_cavity_onoff / set_wake_onoff / setfield
 *           (src/lib/accel/aline.cpp:258-293, :429-463, :980-1031)
 *   accp_track / accp_track_device
 *        <- ALine::pass(double*, int, unsigned st, unsigned ed),
 *           ALine::pass(double*, int), ALine::pass(double*, int, int nturn)
 *           (src/lib/accel/accel.h:625-627, src/lib/accel/aline.cpp:313-363)
 *   accp_beam_stat
 *        <- calc_beam_stat (src/lib/accel/utility.cpp:322-373), as used by
 *           AMonitor::pass (src/lib/accel/accel.cpp:437-475)
 *   accp_batch_* / accp_track_batched
 *        <- no reference counterpart (the reference tracks one lattice per
 *           process); each seed behaves exactly as accp_track on that seed.
 *   accp_adjoint_*
 *        <- replaces the finite-difference responses of ALine::calctuneresp /
 *           calcchromresp (src/lib/accel/aline.cpp:1187-1242, :1310-1355).
 */
#ifndef ACCP_H
#define ACCP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACCP_ABI_VERSION 3

/* ---- error codes ------------------------------------------------------- */
#define ACCP_OK 0
#define ACCP_E_ARG -1     /* invalid argument (null pointer, bad size, range) */
#define ACCP_E_HIP -2     /* HIP runtime failure (no device, launch, copy)     */
#define ACCP_E_NOMEM -3   /* device or host allocation failed                   */
#define ACCP_E_LATTICE -4 /* unsupported element type / inconsistent lattice   */
#define ACCP_E_STATE -5   /* call out of order (e.g. backward before forward)  */

/* ---- element type codes (one per reference element class on the path) --- */
#define ACCP_DRIFT 1     /* ADrift      src/lib/accel/accel.cpp:158-177 */
#define ACCP_QUAD 2      /* AQuad       src/lib/accel/accel.cpp:193-209 */
#define ACCP_CORRECTOR 3 /* ACorrector  src/lib/accel/accel.cpp:270-292 */
#define ACCP_APERTURE 4  /* AAperture   src/lib/accel/accel.cpp:337-387 */
#define ACCP_MONITOR 5   /* AMonitor    src/lib/accel/accel.cpp:437-475 (no-op on coordinates) */
#define ACCP_BEND 6      /* ABend       src/lib/accel/accel.cpp:523-550 */
#define ACCP_BENDMPOLE 7 /* ABendMPole  src/lib/accel/accel.cpp:649-705 */
#define ACCP_STRMPOLE 8  /* AStrMPole   src/lib/accel/accel.cpp:800-847 */
#define ACCP_CAVITY 9    /* ACavity     src/lib/accel/accel.cpp:903-923 */
#define ACCP_LONGWAKE 10 /* longitudinal resonator wake: new, no reference class */
#define ACCP_MATRIX 11   /* AMatrix     src/lib/accel/accel.cpp:1661-1730 */
#define ACCP_SYNCRAD 12  /* ASyncRad    src/lib/accel/accel.cpp:1733-1854 (lumped radiation) */
#define ACCP_CRABCAV 13  /* ACrabCav    src/lib/accel/accel.cpp:927-1000 */
#define ACCP_RFMULT 14   /* ARFMult     src/lib/accel/accel.cpp:1043-1202 (RF multipole) */
#define ACCP_IDEALWIG 15 /* AIdealWig   src/lib/accel/accel.cpp:1386-1537 (ideal wiggler) */
#define ACCP_TRANSWAKE 16 /* ATransWake src/lib/accel/accel.cpp:1546-1657, over the whole beam */
#define ACCP_WIGTABLE 17 /* AWigTable   src/lib/accel/accel.cpp:1205-1382 (kick-table wiggler) */

#define ACCP_MAX_MPOLE 10 /* MAX_MPOLE, src/lib/accel/accel.h:48 */
#define ACCP_WIGTAB_MAX_X 101 /* MAX_WIGTAB_X, src/lib/accel/accel.h:401 */
#define ACCP_WIGTAB_MAX_Y 21  /* MAX_WIGTAB_Y, src/lib/accel/accel.h:402 */
#define ACCP_NAME_LEN 32

/* Radiation flag values of ALine::set_rad_onoff (src/lib/accel/accel.h:577-582,
 * src/lib/accel/aline.cpp:258-293).  A BendMPole radiates with 1 and 2; a
 * SyncRad element acts with 3 (damping and excitation), 4 (damping only) and
 * 5 (excitation only); any other value leaves it (or the bend) off. */
#define ACCP_RAD_OFF 0
#define ACCP_RAD_DAMP 1
#define ACCP_RAD_QE 2
#define ACCP_RAD_LUMP 3
#define ACCP_RAD_LUMP_DAMP 4
#define ACCP_RAD_LUMP_QE 5

/*
 * One element, in the units of the reference lattice file (metres, radians,
 * MV, Hz, MeV).  Fields that a type does not use are ignored.  A record is
 * what AccelElement::create parses for that type plus the state the reference
 * keeps outside the file (nslice, energy, radiation and on/off flags).
 */
typedef struct accp_elem {
  int32_t type;      /* ACCP_* type code                                      */
  int32_t nslice;    /* BendMPole/StrMPole slices (reference default 10)      */
  int32_t max_npole; /* BendMPole/StrMPole MaxNPole (1..10)                   */
  int32_t rad;       /* BendMPole / IdealWig / SyncRad radiation flag ACCP_RAD_* */
  int32_t on;        /* Cavity / CrabCav flag_onoff, LongWake / TransWake on  */
  int32_t nbins;     /* LongWake: number of longitudinal bins                 */
  double len;        /* length [m]                                            */
  double k;          /* Quad K [m^-2]; Bend gradient K                        */
  double theta;      /* Bend/BendMPole bending angle [rad]                    */
  double e1, e2;     /* Bend/BendMPole entrance/exit edge angles [rad]        */
  double energy;     /* BendMPole/Cavity/LongWake beam energy [MeV]           */
  double volt;       /* Cavity voltage [MV]                                   */
  double freq;       /* Cavity frequency [Hz]                                 */
  double phis;       /* Cavity synchronous phase [rad]                        */
  double kickx, kicky;   /* Corrector kicks [rad]                             */
  double xlim[2], ylim[2]; /* Aperture limits [m]                             */
  double polyB[ACCP_MAX_MPOLE]; /* normal multipoles, B[n] in m^-(n+1)       */
  double polyA[ACCP_MAX_MPOLE]; /* skew multipoles                           */
  double wake_rs;    /* LongWake shunt impedance [Ohm]                        */
  double wake_q;     /* LongWake quality factor                               */
  double wake_fr;    /* LongWake resonant frequency [Hz]                      */
  double wake_charge;/* LongWake bunch charge [C]; TransWake Ne [nC]          */
  double wake_zmin, wake_zmax; /* LongWake binning window in z [m]            */
  double mat[36];    /* Matrix: row-major 6x6 map                             */
  double u0, jx, jy; /* SyncRad: energy loss per turn (units of energy), damping partitions */
  double emitx, emity, sigdpp, betax, betay; /* SyncRad equilibrium beam      */
  double devoltx, devolty; /* CrabCav deflecting voltages (units of energy)  */
  double sigphi, sigdvv;   /* CrabCav / RFMult phase and amplitude noise      */
  double wig_lambda, wig_h; /* IdealWig period [m] and peak curvature [1/m]  */
  double tw_w[2], tw_fr[2], tw_q[2]; /* TransWake x/y: strength W, resonant
                                      * frequency [GHz], quality factor; the
                                      * bunch charge [nC] is wake_charge    */
  int32_t wig_type;         /* IdealWig: 0 horizontal, 1 vertical wiggler     */
  int32_t wt_mx, wt_my;     /* WigTable: grid points in x (2..101) and y (2..21) */
  int32_t reserved;
  const double *wt_table;   /* WigTable: xl[mx], yl[my], then the x kicks and
                             * the y kicks [rad], each my rows of mx values
                             * (row iy, column ix), as in the lattice file.
                             * Read when a lattice is built or updated (the
                             * library keeps its own copy); in the records
                             * accp_lattice_get returns it points at that
                             * copy, valid until the next update or free.   */
  char famname[ACCP_NAME_LEN]; /* family name (host-side lookups only)       */
} accp_elem;

typedef struct accp_lattice accp_lattice; /* device-resident element table   */
typedef struct accp_batch accp_batch;     /* many lattices, one element count */
typedef struct accp_adjoint accp_adjoint; /* reverse-mode turn map context    */

/* Element-table staging inside the tracking kernel. */
#define ACCP_STAGE_AUTO 0   /* library default                                */
#define ACCP_STAGE_SCALAR 1 /* wave-uniform scalar loads through the K$       */
#define ACCP_STAGE_LDS 2    /* whole table copied to LDS per workgroup        */
/* The lattice's element sequence (the call's element range) compiled at
 * run time (hiprtc) into a kernel of its own, with the constants in LDS when
 * they fit: no per-element dispatch in the turn, a ring of identical cells a
 * loop over one cell.  Any lattice (up to 256 elements outside a repeated
 * cell); in a batch, seeds of one structure; otherwise ACCP_E_ARG.  Results
 * differ from the table-walking kernel's in the last bits (the compiler may
 * fuse operations across elements).  The first call for a structure, range
 * and launch shape compiles for about a second; the kernel is kept for the
 * life of the process.  Not for accp_track_device_noalloc (loading a kernel
 * allocates device memory).  Never chosen by ACCP_STAGE_AUTO;
 * accp_autotune tries it. */
#define ACCP_STAGE_SPECIALIZED 3

/* Launch shape of the tracking kernel; zero fields mean "library default".
 * Lost particles are dropped from the working set between chunks of turns
 * (chunks of 1, 2, 4, ... turns up to `compact` turns; 0 = default 256,
 * negative = never compact). */
typedef struct accp_tuning {
  int32_t block;   /* threads per workgroup: 64, 128, 256                     */
  int32_t ppl;     /* particles per lane: 1 or 2                              */
  int32_t staging; /* ACCP_STAGE_*                                            */
  int32_t compact; /* longest chunk of turns between compactions              */
} accp_tuning;

/* Options of one tracking call. */
typedef struct accp_track_opts {
  uint64_t rng_seed;    /* key of the counter-based RNG (quantum excitation)  */
  uint64_t turn0;       /* turn counter of the first tracked turn             */
  uint64_t pid0;        /* particle id of r[0] (RNG stream = pid0 + i)        */
  int32_t mon_interval; /* >0: record the beam every mon_interval turns at    */
                        /* turn start (AMonitor at element 0), else 0         */
  int32_t mon_max;      /* capacity of mon_out in records                     */
  double *mon_out;      /* host: mon_max records of ACCP_STAT_LEN doubles     */
  int32_t *mon_count;   /* host: number of records written                    */
  int32_t *loss_turn;   /* optional, n entries (host for accp_track, device   */
                        /* for accp_track_device): turn of this call in which */
                        /* the particle was lost, -1 if lost on input, nturn  */
                        /* if it survived                                     */
  accp_tuning tuning;   /* launch shape (zeros = default)                     */
} accp_track_opts;

/* One beam statistics record, the quantities calc_beam_stat returns:
 * [0]=turn, [1]=nlive, [2..7]=mean of x..delta, [8..13]=rms of x..delta,
 * [14..19]=<xx'>,<yy'>,<x delta>,<x' delta>,<y z>,<y' z> (central moments). */
#define ACCP_STAT_LEN 20

/* ---- library ----------------------------------------------------------- */
const char *accp_last_error(void);
int accp_abi_version(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int accp_device_count(int *count);

/* ---- lattice ----------------------------------------------------------- */
int accp_lattice_build(const accp_elem *elems, int n, accp_lattice **out);
/* Same topology, new parameters or flags (set_flags / modify_element). */
int accp_lattice_update(accp_lattice *lat, const accp_elem *elems, int n);
int accp_lattice_nelem(const accp_lattice *lat, int *n);
int accp_lattice_free(accp_lattice *lat);
/* No counterpart in the reference (whose ALine has one OpenMP loop).  The
 * launch shape of every tracking call on this lattice that leaves its tuning
 * at zero (ACCP_STAGE_AUTO): accp_track / accp_track_device without a
 * tuning, and the lattice services that track internally (getm66, tunes,
 * responses and corrections, trackDA, trackMA, Twiss, the tune diagnostics,
 * the adjoint's forward pass).  NULL restores the library default.  A
 * caller that wants those services on the specialized kernel sets
 * ACCP_STAGE_SPECIALIZED here.  accp_track_device_noalloc ignores it. */
int accp_lattice_set_tuning(accp_lattice *lat, const accp_tuning *tuning);

/* ---- tracking ---------------------------------------------------------- */
/* ALine::pass over elements [st, ed) repeated nturn times on a host AoS
 * buffer; synchronous.  ed < 0 means "to the end of the line". */
int accp_track(accp_lattice *lat, double *r, int n, int nturn, int st, int ed,
               const accp_track_opts *opts);
/* Same, on a device-resident AoS buffer, enqueued on `stream` (a
 * hipStream_t, NULL = default stream).  Monitor records need the host
 * variant. */
int accp_track_device(accp_lattice *lat, double *d_r, int n, int nturn, int st,
                      int ed, const accp_track_opts *opts, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): accp_track_device with the run's working set (live
 * particle index lists, counts, tile list) in a fixed array of the library
 * (about 8.3 MB on each GPU it runs on, held until the process ends)
 * instead of buffers the library allocates; it allocates nothing.  At most
 * ACCP_NOALLOC_MAX_PARTICLES particles (more: ACCP_E_ARG); no monitor
 * records; no TransWake in the range (its sort needs buffers sized to the
 * beam: ACCP_E_LATTICE).  Results equal accp_track_device's bit for bit.
 * All calls on one GPU share the array: calls that overlap on one GPU race. */
#define ACCP_NOALLOC_MAX_PARTICLES (1 << 20)
int accp_track_device_noalloc(accp_lattice *lat, double *d_r, int n, int nturn, int st,
                              int ed, const accp_track_opts *opts, void *stream);

/* calc_beam_stat on a host AoS buffer, computed on the device. */
int accp_beam_stat(const double *r, int n, double *stat /* ACCP_STAT_LEN */);

/* ---- batched multi-seed tracking ---------------------------------------- */
/* All seeds must have the same number of elements (error seeds of one
 * lattice, or one ring started at different elements) and no wake element. */
int accp_batch_create(accp_lattice *const *seeds, int nseed, accp_batch **out);
/* Track all seeds in one launch sequence.  d_r is one device AoS buffer;
 * particles of seed s are [offsets[s], offsets[s+1]) (offsets: host array of
 * nseed+1).  d_loss_turn: optional device int32[n], as in accp_track_opts. */
int accp_track_batched(accp_batch *b, double *d_r, const int64_t *offsets,
                       int nturn, const accp_tuning *tuning,
                       int32_t *d_loss_turn, void *stream);
/* Capture accp_track_batched into a HIP graph (replay = one launch). */
int accp_batch_capture(accp_batch *b, double *d_r, const int64_t *offsets,
                       int nturn, const accp_tuning *tuning,
                       int32_t *d_loss_turn);
int accp_batch_replay(accp_batch *b, void *stream);
/* Time candidate launch shapes on a short run and return the fastest.  The
 * candidates are tried from the smallest workgroup up; a later one replaces
 * the best only when it is faster by more than 1%. */
int accp_autotune(accp_batch *b, double *d_r, const int64_t *offsets,
                  int nturn_probe, accp_tuning *best, double *best_ms);
int accp_batch_free(accp_batch *b);

/* ---- reverse-mode adjoint of the turn map -------------------------------- */
#define ACCP_FIELD_K 1     /* Quad K                                          */
#define ACCP_FIELD_POLYB 2 /* StrMPole PolyB[index]                           */
#define ACCP_FIELD_POLYA 3 /* StrMPole PolyA[index]                           */

/* One differentiable knob: a family of elements whose field moves together
 * (a shift of the knob adds to every member, as calctuneresp does). */
typedef struct accp_param {
  int32_t field;        /* ACCP_FIELD_*                                       */
  int32_t index;        /* multipole index for POLYB/POLYA                    */
  int32_t nelem;        /* number of member elements                          */
  const int32_t *elems; /* member element indices                             */
} accp_param;

/* The reverse sweep carries every map that is deterministic and acts on one
 * particle: Drift, Quad, Bend, BendMPole (radiation off or classical),
 * StrMPole, Matrix, Cavity, Aperture, Corrector, Monitor, SyncRad with
 * damping only, CrabCav and RFMult without noise, IdealWig (with or without
 * its classical loss) and WigTable.  A lattice holding quantum excitation,
 * RF noise or a wake fails with ACCP_E_LATTICE, at create and again at
 * forward and backward (the lattice can change in between). */
int accp_adjoint_create(accp_lattice *lat, int nparam, const accp_param *params,
                        accp_adjoint **out);
/* Forward: track r0 (host AoS, n particles) for nturn turns, keeping a
 * checkpoint every ckpt_every turns; r_out receives the final coordinates. */
int accp_adjoint_forward(accp_adjoint *adj, const double *r0, int n, int nturn,
                         int ckpt_every, double *r_out);
/* Backward: given dL/dr_out (host AoS), return dL/dparam[nparam] and,
 * if dl_dr0 is not NULL, dL/dr0 (host AoS). */
int accp_adjoint_backward(accp_adjoint *adj, const double *dl_dr_out,
                          double *dl_dparam, double *dl_dr0);
/* How the backward pass walks the turns.  SEQUENTIAL: one GPU lane per
 * particle carries dL/d(state) back through every turn.  SEGMENT_JACOBIAN:
 * the transposed Jacobian of every checkpoint segment (and its knob
 * derivatives) is built by 6 lanes per particle and segment, all segments
 * at once, then chained: six times the arithmetic on many more lanes, for
 * beams too small to fill the GPU.  AUTO (default) picks by lane count.
 * Both give the same gradient up to round-off. */
#define ACCP_REVERSE_AUTO 0
#define ACCP_REVERSE_SEQUENTIAL 1
#define ACCP_REVERSE_SEGMENT_JACOBIAN 2
int accp_adjoint_set_reverse(accp_adjoint *adj, int mode);
int accp_adjoint_free(accp_adjoint *adj);

/* ---- ALine's lattice services on the GPU tracker -------------------------- */
/* These mirror the ALine methods the command-file driver calls
 * (src/lib/accel/accel.h:627-642), with the same arguments and arithmetic;
 * every ray or particle is tracked by accp_track.  They change the lattice
 * the way the reference changes its ALine (family strengths, cavity phase,
 * flags); accp_lattice_get returns the records afterwards. */

/* Copy of the lattice's element records (n = accp_lattice_nelem). */
int accp_lattice_get(const accp_lattice *lat, accp_elem *out, int n);

#define ACCP_DELTA_DIFF 1.0e-8 /* DELTA_DIFF_TRACK, src/lib/accel/accel.h:586 */

/* ALine::getm66(m66, st, ed, dp, delta) (src/lib/accel/aline.cpp:891-927):
 * transfer matrix of elements [st, ed) (ed < 0: to the end) from 12 rays at
 * +-delta around (0,0,0,0,0,dp); m66 row-major, m66[6*i+j] = dr_i/dr0_j. */
int accp_getm66(accp_lattice *lat, int st, int ed, double dp, double delta,
                double m66[36]);
/* ALine::calctune (src/lib/accel/aline.cpp:1032-1043): acos(trace/2)/2pi
 * per plane of getm66. */
int accp_calctune(accp_lattice *lat, double dp, double delta, double tune[2]);
/* ALine::calcchrom (src/lib/accel/aline.cpp:1108-1126): tune difference
 * between dp-1e-5 and dp+1e-5 over 2e-5. */
int accp_calcchrom(accp_lattice *lat, double dp, double delta, double chrom[2]);

#define ACCP_TUNE 1  /* knobs: Quad K, or StrMPole PolyB[1]                  */
#define ACCP_CHROM 2 /* knobs: StrMPole PolyB[2]                             */
#define ACCP_RESP_FD 0      /* the reference's +1e-5 forward difference       */
#define ACCP_RESP_ADJOINT 1 /* reverse-mode adjoint through the getm66 rays   */

/* ALine::calctuneresp / calcchromresp (src/lib/accel/aline.cpp:1187-1242,
 * :1310-1355): resp[2*k + p] = d(tune, or chromaticity, of plane p) / d(knob
 * k) for the two families fam1 and fam2 (element indices; a knob shift moves
 * every member).  ACCP_RESP_ADJOINT differentiates the same getm66-based
 * tune exactly instead of by a forward difference; its lattice must be one
 * accp_adjoint_create accepts. */
int accp_calcresp(accp_lattice *lat, int quantity, int method, const int32_t *fam1,
                  int n1, const int32_t *fam2, int n2, double resp[4]);
/* ALine::correcttune / correctchrom (src/lib/accel/aline.cpp:1243-1309,
 * :1356-1418): one 2x2 Newton step of the two families toward target[2],
 * applied to the lattice; target[] then receives the tunes (chromaticities)
 * reached, as the reference returns them. */
int accp_correct(accp_lattice *lat, int quantity, int method, const int32_t *fam1,
                 int n1, const int32_t *fam2, int n2, double target[2]);
/* ALine::setrfphase() (src/lib/accel/aline.cpp:465-501): one turn of a
 * particle at the origin with classical radiation and the cavities off gives
 * the energy loss U0/E; every cavity's phase becomes asin(U0/V_total).  Like
 * the reference it leaves the cavities off and radiation off.  LongWake
 * elements are off for the probe turn (a lone particle would see the whole
 * bunch charge's own field) and keep their flag afterwards.  *phis = 0 and
 * nothing changes when the lattice has no cavity. */
int accp_setrfphase(accp_lattice *lat, double *phis);
/* ALine::trackDA (src/lib/accel/aline.cpp:607-715): nray (3..51) polar rays
 * of nstep points from xmin to xmax (y scaled by ymax/xmax) at z = 1e-5,
 * tracked nturn turns together; rx, ry (51 entries) receive each ray's first
 * lost point (or its last point), *area the enclosed area (positive-x side
 * at half weight if negweight).  loss_turn (optional, nray*nstep) receives
 * the turn in which each point was lost, nturn if it survived. */
int accp_trackda(accp_lattice *lat, double xmax, double xmin, double ymax, int nturn,
                 int nray, int nstep, int negweight, double *rx, double *ry,
                 double *area, int32_t *loss_turn);
/* ALine::getco4 (src/lib/accel/aline.cpp:838-870): 4D closed orbit at dp,
 * rc = (I - M44)^-1 r1 with M44 the transverse block of getm66 at dp and r1
 * the one-turn image of (0,0,0,0,0,dp); radiation and cavities are switched
 * off, as the reference does.  (The reference's own M44 is not the block of
 * getm66: its m66.submatrix call re-points m66 at the uninitialised m44, so
 * its result is r1 alone; this computes the closed orbit.) */
int accp_getco4(accp_lattice *lat, double dp, double rc[4]);
/* ALine::getco6 (src/lib/accel/aline.cpp:871-889): 6D closed orbit with
 * classical radiation and the cavities on (the flags are left so). */
int accp_getco6(accp_lattice *lat, double rc[6]);
/* ALine::calctwiss (src/lib/accel/aline.cpp:1044-1107): with n = 0 the
 * periodic Twiss functions at the ring start, else carried to elements
 * idx[0..n) (increasing) by the transfer matrices between them; per point
 * tw[7*i..] = {s, psix, psiy, betx, alfx, bety, alfy}. */
int accp_calctwiss(accp_lattice *lat, const int32_t *idx, int n, double dp, double delta,
                   double *tw);
/* ALine::calcchrom2 (src/lib/accel/aline.cpp:1127-1158): c = {chromx,
 * chromy, chromx2, chromy2}, the second-order terms from the least-squares
 * slope of the chromaticity over dpp = 1e-3 (i - 6), i = 0..10. */
int accp_calcchrom2(accp_lattice *lat, double c[4]);
/* ALine::trackMA (src/lib/accel/aline.cpp:781-836): for each location
 * locs[k], the ring started there, 2 (nstep+1) particles at
 * +-(dppmin + j (dppmax-dppmin)/nstep), nturn turns; dpp_plus[k] and
 * dpp_minus[k] receive the largest surviving offsets on each side, *avg the
 * mean of min(dpp_plus, -dpp_minus).  All locations are tracked in one
 * batched launch. */
int accp_trackma(accp_lattice *lat, double dppmax, double dppmin, int nturn, int nstep,
                 const int32_t *locs, int nloc, double *dpp_plus, double *dpp_minus,
                 double *avg);

/* ---- tune diagnostics (NAFF) ---------------------------------------------
 * naff (src/lib/accel/utility.cpp:48-96): the frequency, in tune units
 * (0..1/2), of the strongest spectral line of a signal within [xlow, xhi]
 * (clamped to [0.9/m, 1/2]), refined by a golden-section search of the
 * sin^4-windowed projection to a relative bracket of 3e-8.  On the GPU the
 * spectrum and the search run in one workgroup per signal. */
#define ACCP_NAFF_MAX_LEN 4096 /* longest signal; lengths are powers of two >= 4 */
/* nsig host signals of m samples each, data[s*m + i]; tunes[nsig]. */
int accp_naff(const double *data, int nsig, int m, double xlow, double xhi, double *tunes);
/* ALine::tracktune (src/lib/accel/aline.cpp:1159-1186): one particle from
 * (x0, 0, y0, 0, 0, 0); the tunes of its positions at the start and after
 * each of nt - 1 turns, nt the power of two >= nturn (nt <= 4096). */
int accp_tracktune(accp_lattice *lat, int nturn, double x0, double y0, double xlow, double xhi,
                   double ylow, double yhi, double tune[2]);
/* ALine::trackDiffusion (src/lib/accel/aline.cpp:1421-1484): np particles
 * from (x0[i], 0, y0[i], 0, 0, 0), all tracked in one launch; nux, nuy the
 * tunes of the first half of nt positions (nt the power of two >= nturn,
 * nt/2 <= 4096), dif = max(-4 ln nt, ln(2 |dnu| / nt)) with dnu the tune
 * change to the second half; a particle lost within nt - 1 turns gets NaN
 * and adds -3 to *total (the sum of dif, optional). */
int accp_trackdiffusion(accp_lattice *lat, int nturn, int np, const double *x0, const double *y0,
                        double *nux, double *nuy, double *dif, double xlow, double xhi, double ylow,
                        double yhi, double *total);
/* ALine::trackdnudexy (src/lib/accel/aline.cpp:716-779): amplitude-dependent
 * tune shift.  Six particles at x = 1e-5 + k xm/5 (y = 1e-5 + 0.02 x) and six
 * at y = 1e-5 + k ym/5 (x = 1e-5 + 0.02 y), nt turns each (one launch), tunes
 * within 0.1 of the ring's linear tunes; dnudexy = least-squares slopes
 * {dnux/dex, dnuy/dex, dnux/dey, dnuy/dey} against ex = x^2/betx, ey =
 * y^2/bety (betx, bety at the ring start).  rows (optional, 36 doubles)
 * receives {e, nux, nuy} of the twelve particles, the lines the reference
 * prints. */
int accp_trackdnudexy(accp_lattice *lat, int nturn, double xm, double ym, double dnudexy[4],
                      double *rows);

#ifdef __cplusplus
}
#endif
#endif /* ACCP_H */
