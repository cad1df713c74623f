// accp_kernels.h -- launch interface between the C ABI layer and the
// gfx950 kernels (accp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "accp_table.h"

namespace accp {

// One workgroup's share of a batched launch: particles [start, start+count)
// of the concatenated beam, tracked through lattice `lat`.
struct TileDesc {
  int32_t lat;
  int32_t pad;
  int64_t start;
  int64_t count;
};

struct KernelCtx {
  const TileDesc *tiles;     // null: one lattice, tile = blockIdx.x * BLOCK*PPL
  const Hdr *hdr;            // headers of all lattices
  const double *par;         // constants of all lattices
  const int32_t *lat_hdr0;   // first header of lattice s
  const int32_t *lat_par0;   // first constant of lattice s
  const int32_t *lat_npar;   // number of constants of lattice s
  int32_t nelem;             // elements per lattice (one topology)
  int32_t nturn, st, ed;
  double *r;                 // AoS beam
  int64_t n;                 // particles in r
  uint64_t rng_seed, turn0, pid0;
  double *snap;              // monitor snapshots (plain calls only) or null
  int32_t mon_interval, nsnap;
};

hipError_t launch_track(const KernelCtx &ctx, int grid, int block, int ppl, bool lds,
                        size_t lds_bytes, hipStream_t s);
hipError_t launch_beam_stat(const double *snap, int64_t n, int nrec, double turn0, int interval,
                            double *out, hipStream_t s);

}  // namespace accp
