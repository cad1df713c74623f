// accp_internal.h -- library-internal services shared between the ABI
// sources (not part of include/accp.h).
#pragma once
#include <stdint.h>

#include "../../include/accp.h"

namespace accp {

// Tracks the n particles of the host AoS buffer r (updated in place) nturn
// turns over the whole line and records the state at the start of every
// turn: *d_rec receives the device array rec[(t*n + i)*6 + c], t < nturn,
// owned by the lattice and valid until its next tracking call.  Particle i
// draws the random stream of particle id pid0 + i.
int track_record(accp_lattice *lat, double *r, int n, int nturn, const double **d_rec, uint64_t pid0 = 0);

// A fixed device workspace for the working set of up to max_n particles:
// two index lists of max_n int32, two int32 counts, (max_n + 63) / 64 tile
// descriptors (16 bytes each), one int32 tile count, two int64 offsets.
struct FixedWorkspace {
  int32_t *idx0, *idx1, *cnt0, *cnt1;
  void *tiles;
  int32_t *ntiles;
  int64_t *offs;
  int64_t max_n;
};

// accp_track_device with its working set in ws: no device allocation.
int track_device_fixed(accp_lattice *lat, double *d_r, int n, int nturn, int st, int ed,
                       const accp_track_opts *opts, void *stream, const FixedWorkspace &ws);

}  // namespace accp
