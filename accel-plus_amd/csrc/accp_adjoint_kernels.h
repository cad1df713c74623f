// accp_adjoint_kernels.h -- launch interface of the reverse sweep.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "accp_table.h"

namespace accp {

constexpr int ADJ_MAX_PARAM = 16;

struct AdjCtx {
  const Hdr *hdr;
  const double *par;
  const int32_t *knob_slot;  // per element: gradient slot or -1
  const int32_t *knob_kind;  // per element: kind | (multipole index << 8)
  int32_t st, ed;            // element range of one turn
  int32_t t_lo, t_hi;        // turns of this launch, walked from t_hi-1 down to t_lo
  int32_t ckpt_every;        // turns per checkpoint segment
  int32_t seg0;              // first segment held in `states`
  int64_t n;                 // particles
  int64_t group_n;           // particles x segments held in `states`
  const double *states;      // [turn in segment][segment][particle][6] turn-start states
  double *r_end;             // in: state at the end of turn t_hi-1; out: at the start of t_lo
  double *lam;               // in/out: dL/d(state), AoS
  double *grad;              // per particle accumulated knob gradient [n][nparam]
  int32_t nparam;
};

hipError_t launch_adj_reverse(const AdjCtx &c, hipStream_t s);

}  // namespace accp
