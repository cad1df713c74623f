// accp_adjoint.hip -- reverse sweep of the turn map on gfx950.
//
// One particle per lane walks the turns backwards.  Inside a turn it walks
// the elements backwards: each element's input state is rebuilt from its
// output with the closed-form inverse map (accp_adjoint.h) while the adjoint
// dL/d(state) and the knob gradients are carried back through the element's
// vector-Jacobian product.  At every turn boundary the state is reset to the
// exact turn-start state recomputed from the checkpoints, so inversion
// round-off never accumulates beyond one turn.
#include <hip/hip_runtime.h>

#include "accp_adjoint.h"
#include "accp_adjoint_kernels.h"

namespace accp {

// BendMPole without radiation, PolyB[0..MO] in registers.  The record holds
// PolyB[0..mo] (mo = MaxNPole - 1 <= MO); the orders above it are zeros,
// which add exact zeros to the Horner sums.
template <int NV, int MO, bool PLAIN, class PT>
__device__ __forceinline__ void bendmp_rev(Part &p, Adj (&l)[NV], const Hdr h, const PT *P) {
  const int ns = h.a, mo = h.b & 0xff;
  const double psi1 = P[BM_PSI1], psi2 = P[BM_PSI2], h0 = P[BM_IRHO];
  const double l1 = P[BM_L1], l2 = P[BM_L2], k1 = P[BM_K1], k2 = P[BM_K2];
  const double hs1 = P[BM_HS1], hq1 = P[BM_HQ1], hs2 = P[BM_HS2], hq2 = P[BM_HQ2];
  const double k1c = P[BM_K1C], k2c = P[BM_K2C];
  double b[MO + 1];
#pragma unroll
  for (int i = 0; i <= MO; ++i) b[i] = i <= mo ? (double)P[BM_B + i] : 0.0;
  const double c4a = hs1 * 0.5, c4b = hs2 * 0.5, lh1 = k1 * h0, lh2 = k2 * h0;
  if (psi2 != 0) edge_rev_n<NV>(p, l, psi2);
  const HxC ca = hx_consts(p.ipd, l1, hs1, hq1), cb = hx_consts(p.ipd, l2, hs2, hq2);
  const double lhd1 = lh1 * p.d, lhd2 = lh2 * p.d;
  for (int m = 0; m < ns; ++m) {
    hxdrift_c_rev_n<NV>(p, l, ca, h0, c4a);
    bend_kick_rev_n<NV, MO, PLAIN>(p, l, b, k1, h0, lhd1, lh1, k1c, k2c);
    hxdrift_c_rev_n<NV>(p, l, cb, h0, c4b);
    bend_kick_rev_n<NV, MO, PLAIN>(p, l, b, k2, h0, lhd2, lh2, k1c, k2c);
    hxdrift_c_rev_n<NV>(p, l, cb, h0, c4b);
    bend_kick_rev_n<NV, MO, PLAIN>(p, l, b, k1, h0, lhd1, lh1, k1c, k2c);
    hxdrift_c_rev_n<NV>(p, l, ca, h0, c4a);
  }
  if (psi1 != 0) edge_rev_n<NV>(p, l, psi1);
}

template <int NV, int HI, bool THICK, class PT, class GA>
__device__ __forceinline__ void strmp_rev(Part &p, Adj (&l)[NV], const Hdr h, const PT *P, const Knob &kn,
                                          const GA &ga) {
  const PT *C = THICK ? P + 4 : P;
  double b[HI + 1], a[HI + 1];
#pragma unroll
  for (int i = 0; i <= HI; ++i) {
    b[i] = C[i];
    a[i] = C[HI + 1 + i];
  }
  if (!THICK) {
    mpole_kick_rev_n<NV, HI>(p, l, b, a, 1.0, kn, ga);
    return;
  }
  const int ns = h.a;
  const double l1 = P[0], l2 = P[1], k1 = P[2], k2 = P[3];
  const double n1 = l1 * p.ipd, n2 = l2 * p.ipd;
  const double z1 = n1 * (0.5 * p.ipd), z2 = n2 * (0.5 * p.ipd);
  for (int m = 0; m < ns; ++m) {
    drift_c_rev_n<NV>(p, l, n1, z1);
    mpole_kick_rev_n<NV, HI>(p, l, b, a, k1, kn, ga);
    drift_c_rev_n<NV>(p, l, n2, z2);
    mpole_kick_rev_n<NV, HI>(p, l, b, a, k2, kn, ga);
    drift_c_rev_n<NV>(p, l, n2, z2);
    mpole_kick_rev_n<NV, HI>(p, l, b, a, k1, kn, ga);
    drift_c_rev_n<NV>(p, l, n1, z1);
  }
}

// StrMPole with no field: drifts, plus the knob derivative if a knob sits on it.
template <int NV, class PT, class GA>
__device__ __forceinline__ void strmp_empty_rev(Part &p, Adj (&l)[NV], const Hdr h, const PT *P, bool thick,
                                                const Knob &kn, const GA &ga) {
  if (!thick) {
    mpole_knob_only_n<NV>(p, l, 1.0, kn, ga);
    return;
  }
  const int ns = h.a;
  const double l1 = P[0], l2 = P[1], k1 = P[2], k2 = P[3];
  const double n1 = l1 * p.ipd, n2 = l2 * p.ipd;
  const double z1 = n1 * (0.5 * p.ipd), z2 = n2 * (0.5 * p.ipd);
  for (int m = 0; m < ns; ++m) {
    drift_c_rev_n<NV>(p, l, n1, z1);
    mpole_knob_only_n<NV>(p, l, k1, kn, ga);
    drift_c_rev_n<NV>(p, l, n2, z2);
    mpole_knob_only_n<NV>(p, l, k2, kn, ga);
    drift_c_rev_n<NV>(p, l, n2, z2);
    mpole_knob_only_n<NV>(p, l, k1, kn, ga);
    drift_c_rev_n<NV>(p, l, n1, z1);
  }
}

template <int NV, bool THICK, class PT, class GA>
__device__ __forceinline__ void strmp_rev_dispatch(Part &p, Adj (&l)[NV], const Hdr h, const PT *P,
                                                   const Knob &kn, const GA &ga) {
  // the record keeps every coefficient up to hi; the STRMP_SINGLE flag only
  // serves the forward kernel
  switch (h.b == -1 ? -1 : (h.b & 0xff)) {
    case -1: strmp_empty_rev<NV>(p, l, h, P, THICK, kn, ga); break;
    case 0: strmp_rev<NV, 0, THICK>(p, l, h, P, kn, ga); break;
    case 1: strmp_rev<NV, 1, THICK>(p, l, h, P, kn, ga); break;
    case 2: strmp_rev<NV, 2, THICK>(p, l, h, P, kn, ga); break;
    case 3: strmp_rev<NV, 3, THICK>(p, l, h, P, kn, ga); break;
    case 4: strmp_rev<NV, 4, THICK>(p, l, h, P, kn, ga); break;
    case 5: strmp_rev<NV, 5, THICK>(p, l, h, P, kn, ga); break;
    case 6: strmp_rev<NV, 6, THICK>(p, l, h, P, kn, ga); break;
    case 7: strmp_rev<NV, 7, THICK>(p, l, h, P, kn, ga); break;
    case 8: strmp_rev<NV, 8, THICK>(p, l, h, P, kn, ga); break;
    default: strmp_rev<NV, 9, THICK>(p, l, h, P, kn, ga); break;
  }
}

// FULL = false leaves out the reverse maps of the trigonometric quad, the
// linear Bend, the Matrix and the elements that move delta inside them
// (radiating BendMPole, SyncRad, CrabCav, RFMult, IdealWig) or read a kick
// table (WigTable), whose registers would halve the occupancy of every
// lattice's sweep; the host picks the variant per lattice.  The host also
// refuses what has no reverse map (noise, quantum excitation, wakes).
//
// NV adjoint vectors at once (l[0..NV-1], knob derivatives through ga) in
// the lean variant; the full variant carries one.
template <bool FULL, int NV, class PT, class GA>
__device__ __forceinline__ void element_rev_body(Part &p, Adj (&l)[NV], const Hdr h, const PT *P,
                                                 const Knob &kn, const GA &ga) {
  static_assert(!FULL || NV == 1, "the full variant carries one adjoint vector");
  switch (h.op) {
    case OP_DRIFT:
      drift_nl_rev_n<NV>(p, l, P[0]);
      break;
    case OP_QUAD_K0:
      drift_kzero_rev_n<NV>(p, l, P[0]);
      break;
    case OP_MATRIX:
      if constexpr (FULL) matrix_rev(p, l[0], P[0], P + 1);
      break;
    case OP_QUAD: {
      const double L = P[0], K = P[1], aK = P[2], L2 = P[3];
      const bool foc = K > 0;
      const double sK = foc ? 1.0 : -1.0;
      switch (h.a) {
        case 0:
          if constexpr (FULL) quad_trig_rev(p, l[0], L, L2, foc, aK, sK, kn, ga.g);
          break;
        case 4: quad_series_rev_n<NV, 4>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        case 6: quad_series_rev_n<NV, 6>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        case 8: quad_series_rev_n<NV, 8>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        case 10: quad_series_rev_n<NV, 10>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        case 12: quad_series_rev_n<NV, 12>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        case 14: quad_series_rev_n<NV, 14>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
        default: quad_series_rev_n<NV, 16>(p, l, P + 4, L, L2, foc, aK, sK, kn, ga); break;
      }
      break;
    }
    case OP_BENDMP: {
      // plain: both curvature coefficients zero (always so for MaxNPole = 1)
      const bool plain = (h.b >> 8) & 1;
      switch (h.b & 0xff) {
        case 0: bendmp_rev<NV, 0, true>(p, l, h, P); break;
        case 1: plain ? bendmp_rev<NV, 1, true>(p, l, h, P) : bendmp_rev<NV, 1, false>(p, l, h, P); break;
        case 2: plain ? bendmp_rev<NV, 2, true>(p, l, h, P) : bendmp_rev<NV, 2, false>(p, l, h, P); break;
        case 3: plain ? bendmp_rev<NV, 3, true>(p, l, h, P) : bendmp_rev<NV, 3, false>(p, l, h, P); break;
        default: plain ? bendmp_rev<NV, 9, true>(p, l, h, P) : bendmp_rev<NV, 9, false>(p, l, h, P); break;
      }
      break;
    }
    case OP_BEND: {  // edge, bend6, edge (ABend::pass, accel.cpp:523-550)
      if constexpr (FULL) {
        const double psi1 = P[4], psi2 = P[5];
        if (psi2 != 0) edge_rev(p, l[0], psi2);
        bend_linear_rev(p, l[0], P[0], P[1], P[2]);
        if (psi1 != 0) edge_rev(p, l[0], psi1);
      }
      break;
    }
    case OP_STRMP:
      strmp_rev_dispatch<NV, true>(p, l, h, P, kn, ga);
      break;
    case OP_STRMP_THIN:
      strmp_rev_dispatch<NV, false>(p, l, h, P, kn, ga);
      break;
    case OP_CAVITY:
      cavity_rev_n<NV>(p, l, P[0], P[1], P[2], P[3], h.a != 0);
      break;
    case OP_APERTURE:
      // a particle that reaches the end alive passed every aperture: the
      // element acts on it as its drift (or the identity for len = 0)
      if (P[0] != 0) drift_nl_rev_n<NV>(p, l, P[0]);
      break;
    case OP_CORRECTOR:
      corrector_rev_n<NV>(p, l, P[0], P[1], P[2]);
      break;
    case OP_BENDMP_RAD:
      if constexpr (FULL) bendmp_rad_rev(p, l[0], h, P);
      break;
    case OP_SYNCRAD:
      if constexpr (FULL) syncrad_rev(p, l[0], P, (h.a & 1) != 0);
      break;
    case OP_CRABCAV:
      if constexpr (FULL) crabcav_rev(p, l[0], P, (h.a & 1) != 0);
      break;
    case OP_RFMULT:
      if constexpr (FULL) rfmult_rev(p, l[0], P, h.b);
      break;
    case OP_IDEALWIG:
      if constexpr (FULL) idealwig_rev(p, l[0], P, h.a, h.b & 0xff, ((h.b >> 8) & 1) != 0);
      break;
    case OP_WIGTABLE:
      if constexpr (FULL) wigtable_rev(p, l[0], P, h.a, h.b & 0xffff, h.b >> 16);
      break;
    default:
      break;
  }
}

// The knob derivatives of one element (all to its one slot) summed in
// registers and added to the gradient once after the element.
template <int NV>
struct ElemAcc {
  double *a;
  __device__ void operator()(int v, int, double x) const { a[v] += x; }
};
template <bool FULL, int NV, class PT, class GA>
__device__ __forceinline__ void element_rev(Part &p, Adj (&l)[NV], const Hdr h, const PT *P, const Knob &kn,
                                            const GA &ga) {
  if constexpr (FULL) {
    element_rev_body<FULL, NV>(p, l, h, P, kn, ga);
  } else {
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    element_rev_body<FULL, NV>(p, l, h, P, kn, ElemAcc<NV>{acc});
    if (kn.kind != 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) ga(v, kn.slot, acc[v]);
    }
  }
}

template <bool LDS>
__device__ __forceinline__ void load_rev(const Hdr *hdr, const int2 *knob, int e, Hdr &h, Knob &kn) {
  Hdr v = hdr[e];
  int2 k = knob[e];
  if (LDS) {
    v.op = __builtin_amdgcn_readfirstlane(v.op);
    v.off = __builtin_amdgcn_readfirstlane(v.off);
    v.a = __builtin_amdgcn_readfirstlane(v.a);
    v.b = __builtin_amdgcn_readfirstlane(v.b);
    k.x = __builtin_amdgcn_readfirstlane(k.x);
    k.y = __builtin_amdgcn_readfirstlane(k.y);
  }
  h = v;
  kn = Knob{k.x, k.y & 0xff, k.y >> 8};
}

// Element table pointers of a reverse kernel.  With LDS the table (headers,
// knob slot/kind pairs, parameters) is copied into LDS once per workgroup.
struct RevTable {
  const Hdr *hdr;
  const int2 *knob;
  const double *par;
};

template <int BLOCK, bool LDS>
__device__ __forceinline__ RevTable stage_table(const AdjCtx &c, unsigned char *lds_raw) {
  RevTable T{c.hdr, reinterpret_cast<const int2 *>(c.knob), c.par};
  if (LDS) {
    Hdr *lh = reinterpret_cast<Hdr *>(lds_raw);
    int2 *lk = reinterpret_cast<int2 *>(lds_raw + sizeof(Hdr) * c.ed);
    double *lp = reinterpret_cast<double *>(lds_raw + (sizeof(Hdr) + sizeof(int2)) * c.ed);
    for (int e = threadIdx.x; e < c.ed; e += BLOCK) {
      lh[e] = T.hdr[e];
      lk[e] = T.knob[e];
    }
    for (int k = threadIdx.x; k < c.npar; k += BLOCK) lp[k] = T.par[k];
    __syncthreads();
    T = RevTable{lh, lk, lp};
  }
  return T;
}

// Walks turns t_hi-1 .. t_lo of one particle backwards.  `q` indexes the
// particle inside the recomputed group of segments; at the start of every
// turn the state is reset to the exact recomputed one.  The next element's
// header and knob are fetched while the current element runs.
template <bool LDS, bool FULL, int NV, class GA>
__device__ __forceinline__ void reverse_turns(const AdjCtx &c, const RevTable &T, int64_t q, int t_lo,
                                              int t_hi, Part &p, Adj (&l)[NV], const GA &ga) {
  for (int t = t_hi - 1; t >= t_lo; --t) {
    // exact state at the start of turn t, loaded early: nothing below
    // depends on it until the turn is done
    const int j = t % c.ckpt_every;
    const double *s = c.states + ((int64_t)j * c.group_n + q) * 6;
    const double s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3], s4 = s[4], s5 = s[5];
    Hdr h;
    Knob kn;
    load_rev<LDS>(T.hdr, T.knob, c.ed - 1, h, kn);
    for (int e = c.ed - 1; e >= c.st; --e) {
      Hdr hn;
      Knob kn_next;
      load_rev<LDS>(T.hdr, T.knob, e > c.st ? e - 1 : e, hn, kn_next);
      element_rev<FULL, NV>(p, l, h, T.par + h.off, kn, ga);
      h = hn;
      kn = kn_next;
    }
    p.x = s0; p.px = s1; p.y = s2; p.py = s3; p.z = s4;
    set_delta(p, s5);
  }
}

template <int BLOCK, bool LDS, bool FULL>
__global__ __launch_bounds__(BLOCK) void k_adj_reverse(AdjCtx c) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const RevTable T = stage_table<BLOCK, LDS>(c, lds_raw);
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= c.n) return;
  double *pe = c.r_end + 6 * i;
  double *pl = c.lam + 6 * i;
  Part p;
  p.x = pe[0]; p.px = pe[1]; p.y = pe[2]; p.py = pe[3]; p.z = pe[4];
  set_delta(p, pe[5]);
  Adj l{pl[0], pl[1], pl[2], pl[3], pl[4], pl[5]};
  double grad[ADJ_MAX_PARAM];
#pragma unroll
  for (int k = 0; k < ADJ_MAX_PARAM; ++k) grad[k] = 0.0;
  if (p.x != p.x) {
    // lost particles are left out of the gradient
    for (int k = 0; k < 6; ++k) pl[k] = 0.0;
    return;
  }
  // turns t_hi-1 .. t_lo, segment by segment
  for (int t1 = c.t_hi; t1 > c.t_lo;) {
    const int seg = (t1 - 1) / c.ckpt_every;
    const int t0 = seg * c.ckpt_every > c.t_lo ? seg * c.ckpt_every : c.t_lo;
    reverse_turns<LDS, FULL, 1>(c, T, (int64_t)(seg - c.seg0) * c.n + i, t0, t1, p, as_vec(l),
                                GradAcc1{grad});
    t1 = t0;
  }
  pe[0] = p.x; pe[1] = p.px; pe[2] = p.y; pe[3] = p.py; pe[4] = p.z; pe[5] = p.d;
  pl[0] = l.x; pl[1] = l.px; pl[2] = l.y; pl[3] = l.py; pl[4] = l.z; pl[5] = l.d;
  double *g = c.grad + i * c.nparam;
  for (int k = 0; k < c.nparam; ++k) g[k] += grad[k];
}

template <int BLOCK, bool LDS, bool FULL>
__global__ __launch_bounds__(BLOCK) void k_adj_segjac(AdjCtx c) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const RevTable T = stage_table<BLOCK, LDS>(c, lds_raw);
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (gid >= c.group_n * 6) return;
  const int col = (int)(gid % 6);
  const int64_t q = gid / 6;  // segment-in-group * n + particle
  const int seg = c.seg0 + (int)(q / c.n);
  const double *pe = c.r_end + 6 * q;
  Part p;
  p.x = pe[0]; p.px = pe[1]; p.y = pe[2]; p.py = pe[3]; p.z = pe[4];
  set_delta(p, pe[5]);
  // a particle lost by the end of its segment is lost at the end of the run
  // and is never read back
  if (p.x != p.x) return;
  Adj l{col == 0 ? 1.0 : 0.0, col == 1 ? 1.0 : 0.0, col == 2 ? 1.0 : 0.0,
        col == 3 ? 1.0 : 0.0, col == 4 ? 1.0 : 0.0, col == 5 ? 1.0 : 0.0};
  double grad[ADJ_MAX_PARAM];
#pragma unroll
  for (int k = 0; k < ADJ_MAX_PARAM; ++k) grad[k] = 0.0;
  const int t0 = seg * c.ckpt_every;
  const int t1 = t0 + c.ckpt_every < c.nturn ? t0 + c.ckpt_every : c.nturn;
  reverse_turns<LDS, FULL, 1>(c, T, q, t0, t1, p, as_vec(l), GradAcc1{grad});
  double *o = c.jac + (q * 6 + col) * (6 + c.nparam);
  o[0] = l.x; o[1] = l.px; o[2] = l.y; o[3] = l.py; o[4] = l.z; o[5] = l.d;
  for (int k = 0; k < c.nparam; ++k) o[6 + k] = grad[k];
}

// Knob gradients of NV adjoint vectors, G slots each.
template <int NV, int G>
struct GradAccN {
  double (*g)[G];
  __device__ void operator()(int v, int slot, double x) const { g[v][slot] += x; }
};

// Segment Jacobians, lean variant, NV columns per lane: one lane per
// (segment, particle, group of NV columns) carries the adjoints of its
// columns through each element together, so the element's inverse map and
// partial derivatives are formed once for all of them (6 / NV lanes per
// segment and particle instead of 6).  Same output as k_adj_segjac.
template <int BLOCK, bool LDS, int NV, int G>
__global__ __launch_bounds__(BLOCK) void k_adj_segjac_v(AdjCtx c) {
  static_assert(6 % NV == 0, "NV divides the six columns");
  constexpr int NG = 6 / NV;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const RevTable T = stage_table<BLOCK, LDS>(c, lds_raw);
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (gid >= c.group_n * NG) return;
  const int part = (int)(gid % NG);
  const int64_t q = gid / NG;  // segment-in-group * n + particle
  const int seg = c.seg0 + (int)(q / c.n);
  const double *pe = c.r_end + 6 * q;
  Part p;
  p.x = pe[0]; p.px = pe[1]; p.y = pe[2]; p.py = pe[3]; p.z = pe[4];
  set_delta(p, pe[5]);
  if (p.x != p.x) return;
  Adj l[NV];
  double grad[NV][G];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int col = part * NV + v;
    l[v] = Adj{col == 0 ? 1.0 : 0.0, col == 1 ? 1.0 : 0.0, col == 2 ? 1.0 : 0.0,
               col == 3 ? 1.0 : 0.0, col == 4 ? 1.0 : 0.0, col == 5 ? 1.0 : 0.0};
#pragma unroll
    for (int k = 0; k < G; ++k) grad[v][k] = 0.0;
  }
  const int t0 = seg * c.ckpt_every;
  const int t1 = t0 + c.ckpt_every < c.nturn ? t0 + c.ckpt_every : c.nturn;
  reverse_turns<LDS, false, NV>(c, T, q, t0, t1, p, l, GradAccN<NV, G>{grad});
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double *o = c.jac + (q * 6 + part * NV + v) * (6 + c.nparam);
    o[0] = l[v].x; o[1] = l[v].px; o[2] = l[v].y; o[3] = l[v].py; o[4] = l[v].z; o[5] = l[v].d;
    for (int k = 0; k < c.nparam; ++k) o[6 + k] = grad[v][k];
  }
}

__global__ __launch_bounds__(64) void k_adj_combine(AdjCtx c) {
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (i >= c.n) return;
  double *pl = c.lam + 6 * i;
  if (c.r_final[6 * i] != c.r_final[6 * i]) {
    for (int k = 0; k < 6; ++k) pl[k] = 0.0;
    return;
  }
  const int w = 6 + c.nparam;
  double lam[6] = {pl[0], pl[1], pl[2], pl[3], pl[4], pl[5]};
  double *g = c.grad + i * c.nparam;
  for (int s = c.nseg_g - 1; s >= 0; --s) {
    const double *J = c.jac + ((int64_t)s * c.n + i) * 6 * w;
    for (int k = 0; k < c.nparam; ++k) {
      double acc = g[k];
      for (int col = 0; col < 6; ++col) acc += lam[col] * J[col * w + 6 + k];
      g[k] = acc;
    }
    double nl[6];
    for (int r = 0; r < 6; ++r) {
      double acc = 0.0;
      for (int col = 0; col < 6; ++col) acc += lam[col] * J[col * w + r];
      nl[r] = acc;
    }
    for (int r = 0; r < 6; ++r) lam[r] = nl[r];
  }
  for (int k = 0; k < 6; ++k) pl[k] = lam[k];
}

hipError_t launch_adj_reverse(const AdjCtx &c, hipStream_t s) {
  constexpr int B = 64;
  const int64_t grid = (c.n + B - 1) / B;
  if (c.lds_bytes > 0 && c.full)
    hipLaunchKernelGGL((k_adj_reverse<B, true, true>), dim3((unsigned)grid), dim3(B), c.lds_bytes, s, c);
  else if (c.lds_bytes > 0)
    hipLaunchKernelGGL((k_adj_reverse<B, true, false>), dim3((unsigned)grid), dim3(B), c.lds_bytes, s, c);
  else if (c.full)
    hipLaunchKernelGGL((k_adj_reverse<B, false, true>), dim3((unsigned)grid), dim3(B), 0, s, c);
  else
    hipLaunchKernelGGL((k_adj_reverse<B, false, false>), dim3((unsigned)grid), dim3(B), 0, s, c);
  return hipGetLastError();
}

// Many lanes here: 256-lane workgroups let four waves share one LDS copy of
// the table (with a 64-lane workgroup a table of tens of KB would cap a CU
// at two or three waves, below what the registers allow).
hipError_t launch_adj_segjac(const AdjCtx &c, hipStream_t s) {
  if (!c.full && c.nparam <= ADJ_SEGJAC_VEC_G) {
    constexpr int NV = ADJ_SEGJAC_VEC_NV, G = ADJ_SEGJAC_VEC_G;
    if (c.lds_bytes > 0) {
      constexpr int B = 256;
      const int64_t grid = (c.group_n * (6 / NV) + B - 1) / B;
      hipLaunchKernelGGL((k_adj_segjac_v<B, true, NV, G>), dim3((unsigned)grid), dim3(B), c.lds_bytes, s, c);
    } else {
      constexpr int B = 64;
      const int64_t grid = (c.group_n * (6 / NV) + B - 1) / B;
      hipLaunchKernelGGL((k_adj_segjac_v<B, false, NV, G>), dim3((unsigned)grid), dim3(B), 0, s, c);
    }
    return hipGetLastError();
  }
  if (c.lds_bytes > 0) {
    constexpr int B = 256;
    const int64_t grid = (c.group_n * 6 + B - 1) / B;
    if (c.full)
      hipLaunchKernelGGL((k_adj_segjac<B, true, true>), dim3((unsigned)grid), dim3(B), c.lds_bytes, s, c);
    else
      hipLaunchKernelGGL((k_adj_segjac<B, true, false>), dim3((unsigned)grid), dim3(B), c.lds_bytes, s, c);
  } else {
    constexpr int B = 64;
    const int64_t grid = (c.group_n * 6 + B - 1) / B;
    if (c.full)
      hipLaunchKernelGGL((k_adj_segjac<B, false, true>), dim3((unsigned)grid), dim3(B), 0, s, c);
    else
      hipLaunchKernelGGL((k_adj_segjac<B, false, false>), dim3((unsigned)grid), dim3(B), 0, s, c);
  }
  return hipGetLastError();
}

hipError_t launch_adj_combine(const AdjCtx &c, hipStream_t s) {
  const int64_t grid = (c.n + 63) / 64;
  hipLaunchKernelGGL(k_adj_combine, dim3((unsigned)grid), dim3(64), 0, s, c);
  return hipGetLastError();
}

}  // namespace accp
