// accp_ctx.h -- what one tracking launch receives: the tile list over the
// working set, the element tables of every lattice and the run's options.
// Shared by the prebuilt kernels and the lattice-specialized tracker that is
// compiled at run time (so it includes nothing from the HIP runtime).
#pragma once
#include <stdint.h>

#include "accp_table.h"

namespace accp {

// One workgroup's share of a launch: positions [start, start+count) of the
// working-set index list (particles idx[start..]), tracked through lattice
// `lat`.
struct TileDesc {
  int32_t lat;
  int32_t count;
  int64_t start;
};

struct KernelCtx {
  const TileDesc *tiles;     // tile list (device)
  const int32_t *ntiles;     // number of valid tiles (device); later blocks exit
  const int32_t *idx;        // working set: particle index of each position
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
  int32_t *loss_turn;        // per particle: turn lost in (relative to the call) or null
  int32_t turn_rel0;         // turn of this launch relative to the call start
  int32_t first_launch;      // 1: the call's first launch (a particle lost on input gets -1)
};

}  // namespace accp
