// accp_adjoint_kernels.h -- launch interface of the reverse sweep.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "accp_table.h"

namespace accp {

constexpr int ADJ_MAX_PARAM = 16;
// Segment Jacobians of lean lattices with at most ADJ_SEGJAC_VEC_G knobs:
// ADJ_SEGJAC_VEC_NV columns per lane (k_adj_segjac_v).  All six in one lane
// form each element's inverse and partial derivatives once per segment and
// particle; on the 342-element, 8-knob lattice (4096 particles, 32 segments
// of 16 turns) the sweep took 108 ms with one column per lane, 77 with two,
// 62 with three and 49 with six, though six hold about 400 registers (one
// wave per SIMD).
constexpr int ADJ_SEGJAC_VEC_NV = 6;
constexpr int ADJ_SEGJAC_VEC_G = 8;

struct AdjCtx {
  const Hdr *hdr;
  const double *par;
  const int32_t *knob;       // per element: {gradient slot or -1, kind | (multipole index << 8)}
  int32_t st, ed;            // element range of one turn
  int32_t t_lo, t_hi;        // turns of this launch, walked from t_hi-1 down to t_lo
  int32_t ckpt_every;        // turns per checkpoint segment
  int32_t seg0;              // first segment held in `states`
  int64_t n;                 // particles
  int64_t group_n;           // particles x segments held in `states`
  const double *states;      // [turn in segment][segment][particle][6] turn-start states
  double *r_end;             // sequential: in state at the end of turn t_hi-1, out at the start of
                             // t_lo; segment Jacobians: end state of each segment of the group
  double *lam;               // in/out: dL/d(state), AoS
  double *grad;              // per particle accumulated knob gradient [n][nparam]
  int32_t nparam;
  int32_t npar;              // doubles in `par`
  int32_t lds_bytes;         // > 0: stage headers, knobs and parameters in LDS
  int32_t full;              // the lattice holds a map only the full sweep carries
  // segment-Jacobian mode
  int32_t nturn;             // turns of the whole forward run
  int32_t nseg_g;            // segments held in `states` (group_n = nseg_g * n)
  double *jac;               // [segment][particle][column][6 + nparam]
  const double *r_final;     // state after the last turn (lost particles are NaN)
};

// LDS bytes the reverse sweep needs to stage a table of nelem headers and
// npar parameters (headers, then knob slot/kind pairs, then parameters).
inline size_t adj_lds_bytes(size_t nelem, size_t npar) {
  return nelem * (sizeof(Hdr) + 2 * sizeof(int32_t)) + npar * sizeof(double);
}

// Sequential sweep: one lane per particle walks turns t_hi-1 .. t_lo.
hipError_t launch_adj_reverse(const AdjCtx &c, hipStream_t s);
// Segment Jacobians: one lane per (segment, particle, column c) walks its
// segment back from the unit adjoint e_c, giving column c of the segment's
// transposed Jacobian and of its knob derivatives; r_end holds each
// segment's end state [segment][particle][6].  Lean lattices with at most
// ADJ_SEGJAC_VEC_G knobs take ADJ_SEGJAC_VEC_NV columns per lane.
hipError_t launch_adj_segjac(const AdjCtx &c, hipStream_t s);
// Chains the segment Jacobians of the group from the last segment to the
// first: lam <- J^T lam, grad += G^T lam, one lane per particle.
hipError_t launch_adj_combine(const AdjCtx &c, hipStream_t s);

}  // namespace accp
