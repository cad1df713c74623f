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

}  // namespace accp
