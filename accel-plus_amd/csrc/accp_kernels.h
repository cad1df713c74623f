// accp_kernels.h -- launch interface between the C ABI layer and the
// gfx950 kernels (accp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "accp_ctx.h"
#include "accp_table.h"

namespace accp {

// offs[0] = 0, offs[1] = n for a one-lattice run, written by a kernel (no
// host copy).
hipError_t launch_set_offsets(int64_t *offs, int64_t n, hipStream_t s);

// Chunked driver helpers (device-resident bookkeeping, no host round trips,
// so a whole run can be captured in a HIP graph).
hipError_t launch_chunk_init(int32_t *idx, int64_t n, int32_t *loss_turn, int32_t nturn,
                             int32_t *cnt, const int64_t *offs, int nseed, hipStream_t s);
hipError_t launch_make_tiles(const int32_t *cnt, const int64_t *offs, int nseed, int per,
                             TileDesc *tiles, int32_t *ntiles, hipStream_t s);
// Loss turns from the beam alone: -1 for a particle lost on input, nturn for
// the others (what a call that tracks no element leaves).
hipError_t launch_loss_init(const double *r, int64_t n, int32_t *loss_turn, int32_t nturn, hipStream_t s);
hipError_t launch_compact(const TileDesc *tiles, const int32_t *ntiles, int max_tiles,
                          const int32_t *idx_in, const double *r, const int64_t *offs,
                          int32_t *cnt_out, int32_t *idx_out, int nseed, hipStream_t s);

// full: use the kernel variant with every element map (radiation, linear
// Bend, trigonometric quads, high-order BendMPole); otherwise the lean one.
hipError_t launch_track(const KernelCtx &ctx, int grid, int block, int ppl, bool lds,
                        size_t lds_bytes, bool full, hipStream_t s);
hipError_t launch_beam_stat(const double *snap, int64_t n, int nrec, double turn0, int interval,
                            double *out, hipStream_t s);

}  // namespace accp
