// accp_track.h -- device code of the persistent FP64 tracker: the element
// dispatch over the maps of accp_physics.h, the turn loop with parking of
// lost particles, monitor snapshots and loss turns (track_tile).
//
// One workgroup owns a tile of particles of one lattice (a seed), each lane
// holds PPL particles whose 6D state (plus 1/(1+delta)) stays in VGPRs for
// all turns; the element table is read either through wave-uniform scalar
// loads or from a per-workgroup LDS copy.  The prebuilt kernels
// (accp_kernels.hip) walk the table; the lattice-specialized tracker
// (accp_spec.cpp) is this same code compiled at run time with the turn's
// element sequence written out.  It therefore includes nothing beyond the
// device headers of this directory.
//
// Arithmetic is FP64 throughout; there is no dense contraction, so no MFMA.
#pragma once
#include "accp_ctx.h"
#include "accp_physics.h"

namespace accp {

// 1/x for x near 1: the hardware estimate refined by two Newton steps
// (within an ulp; the IEEE division sequence around the same steps adds
// scaling and fix-up instructions that values near 1 never need).
__device__ inline double rcp_near1(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// bndthinkickrad (bend_kick_rad, accp_physics.h) with its two reciprocals
// (the velocity norm of B2perp, atpass.cpp:242, and 1/(1+delta) after the
// loss) taken by rcp_near1: both arguments are within a few percent of 1.
struct RcpNear1 {
  __device__ double operator()(double x) const { return rcp_near1(x); }
};

// Two standard normals for (seed, particle, turn, element, pair): two 53-bit
// uniforms in (0,1) from the four Philox words, Box-Muller with both
// branches (the angle through sincospi: no range reduction on [0, 2)).
__device__ inline void gauss_pair(uint64_t seed, uint64_t pid, uint64_t turn, uint32_t eidx, uint32_t pair,
                                  double &g0, double &g1) {
  uint32_t c[4] = {(uint32_t)pid, (uint32_t)turn, eidx, pair};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t k1 = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
  const uint64_t k2 = ((uint64_t)(c[2] >> 5) << 26) | (uint64_t)(c[3] >> 6);
  const double u1 = ((double)k1 + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)k2 + 0.5) * (1.0 / 9007199254740992.0);
  const double rr = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  g0 = rr * cs;
  g1 = rr * sn;
}

// The normal of slice s: pair s/2, cosine branch for even s, sine for odd.
__device__ inline double gauss_kick(uint64_t seed, uint64_t pid, uint64_t turn, uint32_t eidx, uint32_t slice) {
  double g0, g1;
  gauss_pair(seed, pid, turn, eidx, slice >> 1, g0, g1);
  return (slice & 1) ? g1 : g0;
}

// The slices of a radiating ABendMPole (accel.cpp:649-705): the 2nd-order
// integrator with quantum excitation (flag 2, accel.cpp:693-700) or the
// 4th-order one with the classical loss (flag 1).  PLAIN: no curvature
// gradients (bend_kick_rad<true>).  Delta moves at every kick, so the
// pseudo-drift constants of 1/(1+delta) (hx_consts) are formed after each
// kick, once for the pseudo-drifts up to the next one.
template <int PPL, bool PLAIN, class PT>
__device__ __forceinline__ void bendmp_rad_slices(Part (&p)[PPL], const Hdr h, const PT *P,
                                                  const KernelCtx &ctx, const uint64_t (&pid)[PPL],
                                                  uint64_t turn, int eidx) {
  const int ns = h.a, mo = h.b & 0xff;
  const double h0 = P[BM_IRHO], k1c = P[BM_K1C], k2c = P[BM_K2C], crad = P[BM_CRAD];
  const PT *B = P + BM_B;
  if (h.op == OP_BENDMP_QE) {
    const double sl = P[BM_SL], cq = P[BM_CQ];
    const double half = 0.5 * sl, hs = P[BM_HSH], hq = P[BM_HQH];
    // slices 2j and 2j+1 take the two normals of one Philox block, drawn once
    // per pair of slices (a loop over single slices gets the draw if-converted
    // into every slice)
    const double c4 = hs * 0.5;
    HxC c[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) c[k] = hx_consts(p[k].ipd, half, hs, hq);
    for (int m = 0; m < ns; m += 2) {
      double g0[PPL], g1[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) gauss_pair(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, (uint32_t)(m >> 1), g0[k], g1[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], c[k], h0, c4);
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        bend_kick_rad<PLAIN>(p[k], B, mo, sl, h0, k1c, k2c, crad, cq, 2, g0[k], RcpNear1{});
        c[k] = hx_consts(p[k].ipd, half, hs, hq);
      }
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], c[k], h0, c4);
      if (m + 1 < ns) {
#pragma unroll
        for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], c[k], h0, c4);
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          bend_kick_rad<PLAIN>(p[k], B, mo, sl, h0, k1c, k2c, crad, cq, 2, g1[k], RcpNear1{});
          c[k] = hx_consts(p[k].ipd, half, hs, hq);
        }
#pragma unroll
        for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], c[k], h0, c4);
      }
    }
  } else {
    const double l1 = P[BM_L1], l2 = P[BM_L2], k1 = P[BM_K1], k2 = P[BM_K2];
    const double hs1 = P[BM_HS1], hq1 = P[BM_HQ1], hs2 = P[BM_HS2], hq2 = P[BM_HQ2];
    const double c4a = hs1 * 0.5, c4b = hs2 * 0.5;
    // a pseudo-drift of length l with the constants of the particle's delta
    auto drift = [&](double l, double hs, double hq, double c4) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], hx_consts(p[k].ipd, l, hs, hq), h0, c4);
    };
    auto kick = [&](double kl) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick_rad<PLAIN>(p[k], B, mo, kl, h0, k1c, k2c, crad, 0.0, 1, 0.0, RcpNear1{});
    };
    for (int m = 0; m < ns; ++m) {
      drift(l1, hs1, hq1, c4a);
      kick(k1);
      drift(l2, hs2, hq2, c4b);
      kick(k2);
      drift(l2, hs2, hq2, c4b);
      kick(k1);
      drift(l1, hs1, hq1, c4a);
    }
  }
}

// ABendMPole::pass (accel.cpp:649-705): edge, nslice slices of the 4th-order
// integrator (radiation off or classical, flag <= 1) or of the 2nd-order one
// with quantum excitation (flag 2, accel.cpp:693-700), edge.
template <int PPL, class PT>
__device__ __forceinline__ void bendmp(Part (&p)[PPL], const Hdr h, const PT *P,
                                       const KernelCtx &ctx, const uint64_t (&pid)[PPL],
                                       uint64_t turn, int eidx) {
  const int ns = h.a, mo = h.b & 0xff;
  const bool plain = (h.b >> 8) & 1;
  const double psi1 = P[BM_PSI1], psi2 = P[BM_PSI2], h0 = P[BM_IRHO];
  const double k1c = P[BM_K1C], k2c = P[BM_K2C];
  const PT *B = P + BM_B;
  if (h.b & BM_EDGE_IN) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) edge(p[k], psi1);
  }
  if (h.op == OP_BENDMP_QE || h.op == OP_BENDMP_RAD) {
    if (plain) bendmp_rad_slices<PPL, true>(p, h, P, ctx, pid, turn, eidx);
    else bendmp_rad_slices<PPL, false>(p, h, P, ctx, pid, turn, eidx);
  } else {
    const double l1 = P[BM_L1], l2 = P[BM_L2], k1 = P[BM_K1], k2 = P[BM_K2];
    const double hs1 = P[BM_HS1], hq1 = P[BM_HQ1], hs2 = P[BM_HS2], hq2 = P[BM_HQ2];
    for (int m = 0; m < ns; ++m) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift(p[k], l1, h0, hs1, hq1);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick(p[k], B, mo, plain, k1, h0, k1c, k2c);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift(p[k], l2, h0, hs2, hq2);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick(p[k], B, mo, plain, k2, h0, k1c, k2c);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift(p[k], l2, h0, hs2, hq2);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick(p[k], B, mo, plain, k1, h0, k1c, k2c);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift(p[k], l1, h0, hs1, hq1);
    }
  }
  if (h.b & BM_EDGE_OUT) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) edge(p[k], psi2);
  }
}

// ABendMPole::pass without radiation: delta is fixed inside the element, so
// the pseudo-drift and kick constants that depend on it are formed once;
// PolyB[0..MO] live in registers.
template <int PPL, int MO, bool PLAIN, class PT>
__device__ __forceinline__ void bendmp_fixed(Part (&p)[PPL], const Hdr h, const PT *P) {
  const int ns = h.a;
  const double psi1 = P[BM_PSI1], psi2 = P[BM_PSI2], h0 = P[BM_IRHO];
  const double l1 = P[BM_L1], l2 = P[BM_L2], k1 = P[BM_K1], k2 = P[BM_K2];
  const double hs1 = P[BM_HS1], hq1 = P[BM_HQ1], hs2 = P[BM_HS2], hq2 = P[BM_HQ2];
  const double k1c = P[BM_K1C], k2c = P[BM_K2C];
  double b[MO + 1];
#pragma unroll
  for (int i = 0; i <= MO; ++i) b[i] = P[BM_B + i];
  const double c4a = hs1 * 0.5, c4b = hs2 * 0.5;
  const double lh1 = k1 * h0, lh2 = k2 * h0;
  if constexpr (PLAIN && MO == 0) {
    // a pure dipole: py stays fixed through the slices, z is summed in w
    // (hxdrift_w, bend_kick_w)
    const double lhh1 = lh1 * h0, lhh2 = lh2 * h0;
    HxW fa[PPL], fb[PPL];
    double qa[PPL], qb[PPL], wa[PPL], wb[PPL], py[PPL], py2[PPL], w[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      if (h.b & BM_EDGE_IN) edge(p[k], psi1);
      py[k] = p[k].py;
      py2[k] = p[k].py * p[k].py;
      fa[k] = hx_consts_w(p[k].ipd, l1, hs1, hq1, p[k].py, h0);
      fb[k] = hx_consts_w(p[k].ipd, l2, hs2, hq2, p[k].py, h0);
      qa[k] = lh1 * p[k].d - k1 * b[0];
      qb[k] = lh2 * p[k].d - k2 * b[0];
      const double opd2 = 2.0 * (1.0 + p[k].d);
      wa[k] = opd2 * lh1;
      wb[k] = opd2 * lh2;
      w[k] = 0.0;
    }
    auto slice = [&]() {
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_w(p[k], fa[k], py[k], py2[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick_w(p[k], lhh1, qa[k], wa[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_w(p[k], fb[k], py[k], py2[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick_w(p[k], lhh2, qb[k], wb[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_w(p[k], fb[k], py[k], py2[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) bend_kick_w(p[k], lhh1, qa[k], wa[k], w[k]);
#pragma unroll
      for (int k = 0; k < PPL; ++k) hxdrift_w(p[k], fa[k], py[k], py2[k], w[k]);
    };
    if (ns == 10) {  // the reference's default nslice (accel.h:195): straight-line code
#pragma unroll
      for (int m = 0; m < 10; ++m) slice();
    } else {
      for (int m = 0; m < ns; ++m) slice();
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      dipole_close(p[k], w[k]);
      if (h.b & BM_EDGE_OUT) edge(p[k], psi2);
    }
    return;
  }
  HxC ca[PPL], cb[PPL];
  double lhd1[PPL], lhd2[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    if (h.b & BM_EDGE_IN) edge(p[k], psi1);
    ca[k] = hx_consts(p[k].ipd, l1, hs1, hq1);
    cb[k] = hx_consts(p[k].ipd, l2, hs2, hq2);
    lhd1[k] = lh1 * p[k].d;
    lhd2[k] = lh2 * p[k].d;
  }
  for (int m = 0; m < ns; ++m) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], ca[k], h0, c4a);
#pragma unroll
    for (int k = 0; k < PPL; ++k) bend_kick_c<MO, PLAIN>(p[k], b, k1, h0, lhd1[k], lh1, k1c, k2c);
#pragma unroll
    for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], cb[k], h0, c4b);
#pragma unroll
    for (int k = 0; k < PPL; ++k) bend_kick_c<MO, PLAIN>(p[k], b, k2, h0, lhd2[k], lh2, k1c, k2c);
#pragma unroll
    for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], cb[k], h0, c4b);
#pragma unroll
    for (int k = 0; k < PPL; ++k) bend_kick_c<MO, PLAIN>(p[k], b, k1, h0, lhd1[k], lh1, k1c, k2c);
#pragma unroll
    for (int k = 0; k < PPL; ++k) hxdrift_c(p[k], ca[k], h0, c4a);
  }
  if (h.b & BM_EDGE_OUT) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) edge(p[k], psi2);
  }
}

// AStrMPole::pass with PolyB/PolyA[0..HI] in registers; thick (len > 0,
// nslice Forest-Ruth slices, accel.cpp:829-843) or thin (one integrated kick,
// accel.cpp:824-827).  SINGLE: the field is PolyB[HI] alone (a sextupole,
// say), and the kick sum skips the additions of the zero coefficients.
template <int HI, bool SINGLE>
__device__ __forceinline__ void strmp_sum(const double (&b)[HI + 1], const double (&a)[HI + 1], double x,
                                          double y, double &re, double &im) {
  if constexpr (SINGLE) horner_single<HI>(b[HI], x, y, re, im);
  else if constexpr (HI == 9) estrin9(b, a, x, y, re, im);
  else horner_reg<HI>(b, a, x, y, re, im);
}

template <int PPL, int HI, bool THICK, bool SINGLE, class PT>
__device__ __forceinline__ void strmp_fixed(Part (&p)[PPL], const Hdr h, const PT *P) {
  const PT *C = THICK ? P + 4 : P;
  double b[HI + 1], a[HI + 1];
#pragma unroll
  for (int i = 0; i <= HI; ++i) {
    b[i] = (!SINGLE || i == HI) ? (double)C[i] : 0.0;
    a[i] = SINGLE ? 0.0 : (double)C[HI + 1 + i];
  }
  if (!THICK) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      double re, im;
      strmp_sum<HI, SINGLE>(b, a, p[k].x, p[k].y, re, im);
      p[k].px -= re;
      p[k].py += im;
    }
    return;
  }
  const int ns = h.a;
  const double l1 = P[0], l2 = P[1], k1 = P[2], k2 = P[3];
  double n1[PPL], n2[PPL], z1[PPL], z2[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    n1[k] = l1 * p[k].ipd;
    n2[k] = l2 * p[k].ipd;
    z1[k] = n1[k] * (0.5 * p[k].ipd);
    z2[k] = n2[k] * (0.5 * p[k].ipd);
  }
  if constexpr (SINGLE && HI == 2) {
    // a pure sextupole: sext_slices
    const double kb1 = k1 * b[2], kb2 = k2 * b[2], kb1x2 = 2.0 * kb1, kb2x2 = 2.0 * kb2;
#pragma unroll
    for (int k = 0; k < PPL; ++k) sext_slices(p[k], ns, n1[k], z1[k], n2[k], z2[k], kb1, kb1x2, kb2, kb2x2);
    return;
  }
  for (int m = 0; m < ns; ++m) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      double re, im;
      drift_c(p[k], n1[k], z1[k]);
      strmp_sum<HI, SINGLE>(b, a, p[k].x, p[k].y, re, im);
      p[k].px -= k1 * re;
      p[k].py += k1 * im;
      drift_c(p[k], n2[k], z2[k]);
      strmp_sum<HI, SINGLE>(b, a, p[k].x, p[k].y, re, im);
      p[k].px -= k2 * re;
      p[k].py += k2 * im;
      drift_c(p[k], n2[k], z2[k]);
      strmp_sum<HI, SINGLE>(b, a, p[k].x, p[k].y, re, im);
      p[k].px -= k1 * re;
      p[k].py += k1 * im;
      drift_c(p[k], n1[k], z1[k]);
    }
  }
}

// StrMPole with no field: the slices reduce to drifts of total length len.
template <int PPL, class PT>
__device__ __forceinline__ void strmp_empty(Part (&p)[PPL], const Hdr h, const PT *P) {
  const int ns = h.a;
  const double l1 = P[0], l2 = P[1];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    const double n1 = l1 * p[k].ipd, n2 = l2 * p[k].ipd;
    const double z1 = n1 * (0.5 * p[k].ipd), z2 = n2 * (0.5 * p[k].ipd);
    for (int m = 0; m < ns; ++m) {
      drift_c(p[k], n1, z1);
      drift_c(p[k], n2, z2);
      drift_c(p[k], n2, z2);
      drift_c(p[k], n1, z1);
    }
  }
}

// A thin order-9 StrMPole kick with PolyB[0..1] and PolyA[0..1] zero
// (FUSE_POST_LO2): one integrated kick (accel.cpp:824-827), the sum by
// estrin9_lo2.
template <int PPL, class PT>
__device__ __forceinline__ void strmp_thin9_lo2(Part (&p)[PPL], const PT *C) {
  double b[10], a[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    b[i] = i < 2 ? 0.0 : (double)C[i];
    a[i] = i < 2 ? 0.0 : (double)C[10 + i];
  }
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    double re, im;
    estrin9_lo2(b, a, p[k].x, p[k].y, re, im);
    p[k].px -= re;
    p[k].py += im;
  }
}

template <int PPL, bool THICK, class PT>
__device__ __forceinline__ void strmp_dispatch(Part (&p)[PPL], const Hdr h, const PT *P) {
  if (h.b != -1 && (h.b & STRMP_SINGLE)) {
    switch (h.b & 0xff) {
      case 1: strmp_fixed<PPL, 1, THICK, true>(p, h, P); break;
      case 2: strmp_fixed<PPL, 2, THICK, true>(p, h, P); break;
      default: strmp_fixed<PPL, 3, THICK, true>(p, h, P); break;
    }
    return;
  }
  switch (h.b) {
    case -1:
      if (THICK) strmp_empty<PPL>(p, h, P);
      break;
    case 0: strmp_fixed<PPL, 0, THICK, false>(p, h, P); break;
    case 1: strmp_fixed<PPL, 1, THICK, false>(p, h, P); break;
    case 2: strmp_fixed<PPL, 2, THICK, false>(p, h, P); break;
    case 3: strmp_fixed<PPL, 3, THICK, false>(p, h, P); break;
    case 4: strmp_fixed<PPL, 4, THICK, false>(p, h, P); break;
    case 5: strmp_fixed<PPL, 5, THICK, false>(p, h, P); break;
    case 6: strmp_fixed<PPL, 6, THICK, false>(p, h, P); break;
    case 7: strmp_fixed<PPL, 7, THICK, false>(p, h, P); break;
    case 8: strmp_fixed<PPL, 8, THICK, false>(p, h, P); break;
    default: strmp_fixed<PPL, 9, THICK, false>(p, h, P); break;
  }
}

template <int PPL, bool FULL, class PT>
__device__ __forceinline__ void quad_dispatch(Part (&p)[PPL], const Hdr h, const PT *P) {
  const double L = P[0], K = P[1], aK = P[2], L2 = P[3];
  const bool foc = K > 0;
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    switch (h.a) {
      case 4: quad_series<4>(p[k], P + 4, L, L2, foc, aK); break;
      case 6: quad_series<6>(p[k], P + 4, L, L2, foc, aK); break;
      case 8: quad_series<8>(p[k], P + 4, L, L2, foc, aK); break;
      case 10: quad_series<10>(p[k], P + 4, L, L2, foc, aK); break;
      case 12: quad_series<12>(p[k], P + 4, L, L2, foc, aK); break;
      case 14: quad_series<14>(p[k], P + 4, L, L2, foc, aK); break;
      case 16: quad_series<16>(p[k], P + 4, L, L2, foc, aK); break;
      default:
        // outside the series range: only the full kernel carries the
        // trigonometric map (the host never sends such a quad to the lean one)
        if (FULL) quad(p[k], L, K, aK);
        break;
    }
  }
}

// ---- one element on PPL particles -------------------------------------------
template <int PPL, bool FULL, class PT>
__device__ __forceinline__ void element_body(Part (&p)[PPL], const Hdr h, const PT *P,
                                             const KernelCtx &ctx, const uint64_t (&pid)[PPL],
                                             uint64_t turn, int eidx) {
  switch (h.op) {
    case OP_DRIFT: {
      const double L = P[0];
#pragma unroll
      for (int k = 0; k < PPL; ++k) drift_nl(p[k], L * p[k].ipd);
      break;
    }
    case OP_QUAD:
      quad_dispatch<PPL, FULL>(p, h, P);
      break;
    case OP_QUAD_K0: {
      const double L = P[0];
#pragma unroll
      for (int k = 0; k < PPL; ++k) drift_kzero(p[k], L);
      break;
    }
    case OP_BEND: {
      if (!FULL) break;
      const double L = P[0], kx = P[1], grd = P[2], psi1 = P[4], psi2 = P[5];
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        if (psi1 != 0) edge(p[k], psi1);
        bend_linear(p[k], L, kx, grd);
        if (psi2 != 0) edge(p[k], psi2);
      }
      break;
    }
    case OP_BENDMP: {
      const int mo = h.b & 0xff;
      const bool plain = (h.b >> 8) & 1;
      if (plain) {
        switch (mo) {
          case 0: bendmp_fixed<PPL, 0, true>(p, h, P); break;
          case 1: bendmp_fixed<PPL, 1, true>(p, h, P); break;
          case 2: bendmp_fixed<PPL, 2, true>(p, h, P); break;
          case 3: bendmp_fixed<PPL, 3, true>(p, h, P); break;
          default:
            if (FULL) bendmp<PPL>(p, h, P, ctx, pid, turn, eidx);
            break;
        }
      } else {
        switch (mo) {
          case 1: bendmp_fixed<PPL, 1, false>(p, h, P); break;
          case 2: bendmp_fixed<PPL, 2, false>(p, h, P); break;
          case 3: bendmp_fixed<PPL, 3, false>(p, h, P); break;
          default:
            if (FULL) bendmp<PPL>(p, h, P, ctx, pid, turn, eidx);
            break;
        }
      }
      break;
    }
    case OP_BENDMP_RAD:
    case OP_BENDMP_QE:
      if (FULL) bendmp<PPL>(p, h, P, ctx, pid, turn, eidx);
      break;
    case OP_STRMP:
      strmp_dispatch<PPL, true>(p, h, P);
      break;
    case OP_STRMP_THIN:
      strmp_dispatch<PPL, false>(p, h, P);
      break;
    case OP_CAVITY: {
      const double L = P[0], nv = P[1], w = P[2], phis = P[3];
      const bool on = h.a != 0;
#pragma unroll
      for (int k = 0; k < PPL; ++k) cavity(p[k], L, nv, w, phis, on);
      break;
    }
    case OP_APERTURE: {
      const double L = P[0], x0 = P[1], x1 = P[2], y0 = P[3], y1 = P[4];
#pragma unroll
      for (int k = 0; k < PPL; ++k) aperture(p[k], L, x0, x1, y0, y1);
      break;
    }
    case OP_CORRECTOR: {
      const double L = P[0], kx = P[1], ky = P[2];
#pragma unroll
      for (int k = 0; k < PPL; ++k) corrector(p[k], L, kx, ky);
      break;
    }
    case OP_MATRIX: {
      if (!FULL) break;
      const double L = P[0];
#pragma unroll
      for (int k = 0; k < PPL; ++k) matrix_map(p[k], L, P + 1);
      break;
    }
    case OP_SYNCRAD: {
      if (!FULL) break;
      const bool damp = h.a & 1, qe = h.a & 2;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        double g0 = 0, g1 = 0, g2 = 0;
        if (qe) {
          g0 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, 0);
          g1 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, 1);
          g2 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, 2);
        }
        syncrad(p[k], P, damp, qe, g0, g1, g2);
      }
      break;
    }
    case OP_RFMULT: {
      if (!FULL) break;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        const double g = h.a ? gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, 0) : 0.0;
        rfmult(p[k], P, h.b, h.a != 0, g);
      }
      break;
    }
    case OP_IDEALWIG: {
      if (!FULL) break;
#pragma unroll
      for (int k = 0; k < PPL; ++k) idealwig(p[k], P, h.a, h.b & 0xff, (h.b >> 8) & 1);
      break;
    }
    case OP_WIGTABLE: {
      if (!FULL) break;
#pragma unroll
      for (int k = 0; k < PPL; ++k) wigtable(p[k], P, h.a, h.b & 0xffff, h.b >> 16);
      break;
    }
    case OP_CRABCAV: {
      if (!FULL) break;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        // normals in the reference's draw order (accel.cpp:1015-1022): the
        // phase's first when it draws, then x's and y's amplitudes
        double g0 = 0, g1 = 0, g2 = 0;
        const bool ph = (h.a & 1) && (h.a & 2);
        const uint32_t ja = ph ? 1u : 0u;
        if (ph) g0 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, 0);
        if ((h.a & 1) && (h.a & 4)) {
          g1 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, ja);
          g2 = gauss_kick(ctx.rng_seed, pid[k], turn, (uint32_t)eidx, ja + 1);
        }
        crabcav(p[k], P, h.a, g0, g1, g2);
      }
      break;
    }
    default:
      break;
  }
}

// Lost particles (x NaN) stay frozen, as the reference skips them
// (e.g. accel.cpp:166-167).  Every lane computes every element (no
// divergent branch, no per-element save and restore): a lost particle's
// registers turn to garbage, but its frozen state was written to the beam
// buffer when it was lost ("parked", see track_tile), and that is what the
// launch leaves there.  x NaN absorbs every map, so a lost particle stays
// lost.  A record of the grouped table may also carry a drift run before
// the element (FUSE_PRE, its length first) and a thin StrMPole kick after it
// (FUSE_POST): one dispatch, the same maps in the same order.
template <int PPL, bool FULL, class PT>
__device__ __forceinline__ void apply_element(Part (&p)[PPL], const Hdr h, const PT *P,
                                              const KernelCtx &ctx, const uint64_t (&pid)[PPL],
                                              uint64_t turn, int eidx) {
  if constexpr (FULL) {
    // the host groups only lattices of the lean variant
    element_body<PPL, FULL>(p, h, P, ctx, pid, turn, eidx);
  } else {
    const PT *Pm = P;
    if (h.op & FUSE_PRE) {
      const double L = P[0];
#pragma unroll
      for (int k = 0; k < PPL; ++k) drift_nl(p[k], L * p[k].ipd);
      Pm = P + 1;
    }
    // the commonest elements first, each one test: a pure dipole BendMPole,
    // a pure sextupole, an 8-term series quad, then the general dispatch
    const Hdr hm{h.op & FUSE_OP_MASK, h.off, h.a, h.b};
    if (hm.op == OP_BENDMP && (hm.b & 0x1ff) == (1 << 8)) {
      bendmp_fixed<PPL, 0, true>(p, hm, Pm);
    } else if (hm.op == OP_STRMP && hm.b != -1 && (hm.b & (0xff | STRMP_SINGLE)) == (2 | STRMP_SINGLE)) {
      strmp_fixed<PPL, 2, true, true>(p, hm, Pm);
    } else if (hm.op == OP_QUAD && hm.a == 8) {
      const double L = Pm[0], K = Pm[1], aK = Pm[2], L2 = Pm[3];
#pragma unroll
      for (int k = 0; k < PPL; ++k) quad_series<8>(p[k], Pm + 4, L, L2, K > 0, aK);
    } else {
      element_body<PPL, FULL>(p, hm, Pm, ctx, pid, turn, eidx);
    }
    if (h.op & FUSE_POST) {
      const PT *Pk = Pm + ((h.op >> FUSE_POST_DELTA_SHIFT) & FUSE_POST_MAX_DELTA);
      const int pb = ((h.op >> FUSE_POST_HI_SHIFT) & 0xf) | ((h.op & FUSE_POST_SINGLE) ? STRMP_SINGLE : 0);
      if (pb == 9 && (h.op & FUSE_POST_LO2)) strmp_thin9_lo2<PPL>(p, Pk);
      else if (pb == 9) strmp_fixed<PPL, 9, false, false>(p, Hdr{OP_STRMP_THIN, 0, 0, pb}, Pk);
      else strmp_dispatch<PPL, false>(p, Hdr{OP_STRMP_THIN, 0, 0, pb}, Pk);
    }
  }
}

__device__ __forceinline__ void park(double *dst, const Part &p) {
  dst[0] = p.x; dst[1] = p.px; dst[2] = p.y;
  dst[3] = p.py; dst[4] = p.z; dst[5] = p.d;
}

// Monitor record of turn slot: the state in registers, or for a parked
// particle its frozen state from the beam buffer.
template <int PPL>
__device__ __forceinline__ void snapshot(const KernelCtx &ctx, int slot, const Part (&p)[PPL],
                                         const int64_t (&gidx)[PPL], const bool (&inb)[PPL],
                                         const bool (&parked)[PPL]) {
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    const double *src = ctx.r + 6 * gidx[k];
    const double v[6] = {src[0], src[1], src[2], src[3], src[4], src[5]};
    if (inb[k] && slot < ctx.nsnap) {
      double *dst = ctx.snap + ((int64_t)slot * ctx.n + gidx[k]) * 6;
      dst[0] = parked[k] ? v[0] : p[k].x;
      dst[1] = parked[k] ? v[1] : p[k].px;
      dst[2] = parked[k] ? v[2] : p[k].y;
      dst[3] = parked[k] ? v[3] : p[k].py;
      dst[4] = parked[k] ? v[4] : p[k].z;
      dst[5] = parked[k] ? v[5] : p[k].d;
    }
  }
}

// Headers: the next one is read while the current element runs.  From LDS
// it arrives in vector registers and becomes wave-uniform (readfirstlane)
// only after that element, so the read's latency hides behind it.
struct HdrRaw {
  int32_t op, off, a, b;
};

template <bool LDS>
__device__ __forceinline__ HdrRaw load_hdr_raw(const Hdr *hdr, int e) {
  const Hdr v = hdr[e];
  return HdrRaw{v.op, v.off, v.a, v.b};
}

template <bool LDS>
__device__ __forceinline__ Hdr uniform_hdr(const HdrRaw &v) {
  if (!LDS) return Hdr{v.op, v.off, v.a, v.b};
  return Hdr{__builtin_amdgcn_readfirstlane(v.op), __builtin_amdgcn_readfirstlane(v.off),
             __builtin_amdgcn_readfirstlane(v.a), __builtin_amdgcn_readfirstlane(v.b)};
}

// One element (or group) of a turn, then the parking of particles an
// aperture marked lost: their frozen state goes to the beam buffer now.
template <int PPL, bool FULL, class PT>
__device__ __forceinline__ void track_step(Part (&p)[PPL], const Hdr h, const PT *P, const KernelCtx &ctx,
                                           const uint64_t (&pid)[PPL], uint64_t turn, int e,
                                           const int64_t (&gidx)[PPL], bool (&parked)[PPL]) {
  apply_element<PPL, FULL>(p, h, P, ctx, pid, turn, e);
  if ((FULL ? h.op : (h.op & FUSE_OP_MASK)) == OP_APERTURE) {
#pragma unroll
    for (int k = 0; k < PPL; ++k)
      if (!parked[k] && !is_live(p[k])) {
        park(ctx.r + 6 * gidx[k], p[k]);
        parked[k] = true;
      }
  }
}

// The turn as a walk over the element table: elements [ctx.st, ctx.ed) in
// order, each dispatched on its header.  The header of the next element is
// fetched while the current one runs.
struct TableWalker {
  template <int PPL, bool FULL, bool LDS, class PT>
  __device__ static __forceinline__ void turn(Part (&p)[PPL], const Hdr *hdr, const PT *par, const KernelCtx &ctx,
                                              const uint64_t (&pid)[PPL], uint64_t turn,
                                              const int64_t (&gidx)[PPL], bool (&parked)[PPL]) {
    Hdr h = uniform_hdr<LDS>(load_hdr_raw<LDS>(hdr, ctx.st));
    for (int e = ctx.st; e < ctx.ed; ++e) {
      const HdrRaw hn = load_hdr_raw<LDS>(hdr, e + 1 < ctx.ed ? e + 1 : e);
      track_step<PPL, FULL>(p, h, par + h.off, ctx, pid, turn, e, gidx, parked);
      h = uniform_hdr<LDS>(hn);
    }
  }
};

// ---- the persistent tracker ---------------------------------------------------
// The body of a tracking kernel: one tile of one lattice, every turn of the
// launch, the turn itself walked by Walker (TableWalker reads the element
// table; a lattice-specialized walker is compiled from the table's
// structure at run time).  FULL = false is the lean variant for lattices
// whose quads are all in the series range and that have no radiation, no
// linear Bend and no BendMPole above octupole order; the host picks the
// variant per lattice.
template <int BLOCK, int PPL, bool LDS, bool FULL, class Walker>
__device__ __forceinline__ void track_tile(const KernelCtx &ctx) {
  extern __shared__ __align__(16) unsigned char lds_raw[];

  // tile of this workgroup
  if ((int)blockIdx.x >= *ctx.ntiles) return;
  const TileDesc tile = ctx.tiles[blockIdx.x];
  const int lat = __builtin_amdgcn_readfirstlane(tile.lat);
  const int64_t start = tile.start;
  const int count = tile.count;
  const Hdr *__restrict__ ghdr = ctx.hdr + ctx.lat_hdr0[lat];
  const double *__restrict__ gpar = ctx.par + ctx.lat_par0[lat];

  const Hdr *hdr = ghdr;
  const double *par = gpar;
  if (LDS) {
    Hdr *lh = reinterpret_cast<Hdr *>(lds_raw);
    double *lp = reinterpret_cast<double *>(lds_raw + sizeof(Hdr) * ctx.nelem);
    for (int i = threadIdx.x; i < ctx.nelem; i += BLOCK) lh[i] = ghdr[i];
    const int np = ctx.lat_npar[lat];
    for (int i = threadIdx.x; i < np; i += BLOCK) lp[i] = gpar[i];
    __syncthreads();
    hdr = lh;
    par = lp;
  }

  Part p[PPL];
  uint64_t pid[PPL];
  int64_t gidx[PPL];
  bool inb[PPL];
  int lost_at[PPL];
  bool record[PPL];
  // parked: the particle's frozen state is in ctx.r and the launch leaves it
  // there (lost on input, padding lanes, and particles lost in this launch:
  // at an aperture, where the reference marks them, or, for an orbit that
  // overflowed to NaN, at the end of that turn)
  bool parked[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    const int li = k * BLOCK + (int)threadIdx.x;
    inb[k] = li < count;
    gidx[k] = ctx.idx[start + (inb[k] ? li : 0)];
    pid[k] = ctx.pid0 + (uint64_t)gidx[k];
    const double *src = ctx.r + 6 * gidx[k];
    p[k].x = src[0];
    p[k].px = src[1];
    p[k].y = src[2];
    p[k].py = src[3];
    p[k].z = src[4];
    set_delta(p[k], src[5]);
    // a particle already lost when this launch starts was recorded earlier,
    // unless this is the first launch of the call (lost on input: -1)
    record[k] = inb[k] && ctx.loss_turn && (is_live(p[k]) || ctx.first_launch);
    lost_at[k] = 0x7fffffff;
    parked[k] = !inb[k] || !is_live(p[k]);
    if (!inb[k]) p[k].x = __builtin_nan("");  // padding lanes start lost, never written back
  }

  int t = 0;
  for (; t < ctx.nturn; ++t) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      const bool lv = is_live(p[k]);
      any |= lv && inb[k];
      if (!lv && lost_at[k] == 0x7fffffff) lost_at[k] = ctx.turn_rel0 + t - 1;
      if (!lv && !parked[k]) {  // overflowed in the last turn
        park(ctx.r + 6 * gidx[k], p[k]);
        parked[k] = true;
      }
    }
    if (ctx.snap && t % ctx.mon_interval == 0) snapshot<PPL>(ctx, t / ctx.mon_interval, p, gidx, inb, parked);
    if (!__any(any)) break;  // every particle of this wave is lost: frozen from here on
    const uint64_t turn = ctx.turn0 + (uint64_t)t;
    Walker::template turn<PPL, FULL, LDS>(p, hdr, par, ctx, pid, turn, gidx, parked);
  }
  // monitor slots after an early exit hold the frozen state
  if (ctx.snap) {
    for (int tt = t + 1; tt < ctx.nturn; ++tt)
      if (tt % ctx.mon_interval == 0) snapshot<PPL>(ctx, tt / ctx.mon_interval, p, gidx, inb, parked);
  }
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    if (!inb[k]) continue;
    if (!parked[k]) park(ctx.r + 6 * gidx[k], p[k]);
    if (record[k]) {
      int lt = lost_at[k];
      if (lt == 0x7fffffff && !is_live(p[k])) lt = ctx.turn_rel0 + t - 1;  // lost in the last turn run
      if (lt != 0x7fffffff) ctx.loss_turn[gidx[k]] = lt;
    }
  }
}

}  // namespace accp
