// accp_table.h -- the compiled element table shared by the host compiler and
// the gfx950 kernels.
//
// A lattice is compiled into one flat table: a 16-byte header per element
// {op, off, a, b} plus one packed double array of per-element constants.
// Constants that the reference recomputes on every pass but that do not
// depend on the particle (slice lengths, Forest-Ruth kick/drift lengths,
// edge-focusing strength, radiation prefactors) are evaluated once on the
// host, with the reference's operation order, so the kernels only do the
// particle-dependent arithmetic.
#pragma once
#include <stdint.h>

namespace accp {

enum Op : int32_t {
  OP_NOP = 0,     // Monitor, disabled wake
  OP_DRIFT,       // [L]
  OP_QUAD,        // [L, K, |K|, L^2, 3*NT series coefs]   a=NT series terms (0: trig)
  OP_QUAD_K0,     // [L]              (quad6 with K == 0: the reference's drift())
  OP_BEND,        // [L, angle/L, K, irho, psi1, psi2]
  OP_BENDMP,      // BM_* constants then B0..Bmo             a=nslice b=mo | plain<<8
  OP_BENDMP_RAD,  // same layout, classical radiation in the kicks (flag 1)
  OP_BENDMP_QE,   // same layout, radiation + quantum excitation (flag 2, 2nd-order integrator)
  OP_STRMP,       // [L1, L2, K1, K2, B0..Bhi, A0..Ahi]   a=nslice b=hi | STRMP_SINGLE (-1: no field)
  OP_STRMP_THIN,  // [B0..Bhi, A0..Ahi]                     b=hi | STRMP_SINGLE
  OP_CAVITY,      // [L, V/E, 2*pi*f, phis]                 a=on
  OP_APERTURE,    // [L, xmin, xmax, ymin, ymax]
  OP_CORRECTOR,   // [L, kickx, kicky]
  OP_WAKE,        // collective: the host splits the turn here; no-op inside a pass
  OP_MATRIX,      // [L, M00..M55]                          (row-major 6x6)
  OP_SYNCRAD,     // [L, 1-ax, 1-ay, 1-az, sigx sx, sigy sy, sigdpp sz]   a = damping | excitation<<1
  OP_CRABCAV,     // [L, Vx/E, Vy/E, 2 pi f / c, phis, sigphi, sigdVV]    a = on | phase noise<<1 | amplitude noise<<2
  OP_RFMULT,      // [L, 2 pi f / c, phis, sigphi, B0..Bn-1, A0..An-1]     a = phase noise, b = n
  OP_IDEALWIG,    // [L, k^2, h^2, L/nslice, L/(2 nslice), radiation coef] a = nslice, b = type | radiation<<8
  OP_WIGTABLE,    // [L/2, nslice, xl[mx], yl[my], xkick[my][mx], ykick[my][mx]] a = nslice (0: thin), b = mx | my<<16
  OP_COUNT
};

struct Hdr {
  int32_t op;
  int32_t off;  // first constant in the packed double array
  int32_t a;
  int32_t b;
};

// StrMPole header flag: the field is the single normal term PolyB[hi]
// (PolyA zero, PolyB zero below hi), 1 <= hi <= 3.
constexpr int STRMP_SINGLE = 0x100;

// Grouped table (the tracker's): Hdr::op carries the op code in its low
// byte and, above it, a drift run before the element (its length first
// in the record) and a thin StrMPole kick run after it (PolyB[0..hi],
// PolyA[0..hi] at FUSE_POST_DELTA doubles past the element's own constants).
constexpr int32_t FUSE_OP_MASK = 0xff;
constexpr int32_t FUSE_PRE = 1 << 8;
constexpr int32_t FUSE_POST = 1 << 9;
constexpr int FUSE_POST_DELTA_SHIFT = 10;
constexpr int FUSE_POST_MAX_DELTA = 1023;
constexpr int FUSE_POST_HI_SHIFT = 20;
constexpr int32_t FUSE_POST_SINGLE = 1 << 24;
// The post kick is of order 9 with PolyB[0], PolyB[1], PolyA[0], PolyA[1]
// zero (a field-error record of b3..b10, a3..a10): its sum is z^2 times a
// polynomial of degree 7 (estrin9_lo2).
constexpr int32_t FUSE_POST_LO2 = 1 << 25;

// BendMPole header flags beside hi | plain << 8: an entry or exit edge with
// non-zero focusing (psi1, psi2 != 0).
constexpr int BM_EDGE_IN = 1 << 9;
constexpr int BM_EDGE_OUT = 1 << 10;

// Offsets inside an OP_BENDMP* record.  HS* = irho*L and HQ* = (1/4)*HS*L
// of the pseudo-drift lengths L1, L2 and SL/2; K1C/K2C are PolyB[1]/[2] as the
// curvature terms of bndthinkick read them (PolyB[1] whatever MaxNPole,
// PolyB[2] from MaxNPole 3 on); CQ is
// 55*sqrt(3)/72 * uc/E0 of the quantum-excitation sigma.
enum {
  BM_PSI1 = 0, BM_PSI2, BM_IRHO, BM_L1, BM_L2, BM_K1, BM_K2, BM_SL, BM_CRAD, BM_CQ,
  BM_HS1, BM_HQ1, BM_HS2, BM_HQ2, BM_HSH, BM_HQH, BM_K1C, BM_K2C, BM_B
};

// Forest-Ruth / Yoshida 4th-order coefficients (src/lib/accel/accel.h:44-47,
// same decimal literals).
constexpr double DRIFT1 = 0.6756035959798286638;
constexpr double DRIFT2 = -0.1756035959798286639;
constexpr double KICK1 = 1.351207191959657328;
constexpr double KICK2 = -1.702414383919314656;

constexpr double PI = 3.1415926535897932384626;  // accel.h:23
constexpr double TWOPI = 2 * PI;
constexpr double C0 = 2.99792458e8;               // accel.h:29
constexpr double CGAMMA = 8.846056192e-05;        // atpass.cpp:270

}  // namespace accp
