// accp_noalloc.hip -- accp_track_device_noalloc: tracking with no device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The working set of a run (live-particle index lists, per-seed counts,
// tile list, offsets) lives in `scratch`, a fixed array of this library in
// each GPU's memory, sized for ACCP_NOALLOC_MAX_PARTICLES particles (about
// 8.3 MB per GPU, held until the process ends).  Its device address is taken
// once per device, with that device current, on the first call there, and
// every later call on that device uses the kept address.  All calls on one
// GPU share the array: calls that overlap on one GPU race.  Every run
// rewrites the parts it reads (index lists, counts, tiles and offsets are
// set by kernels at its start), so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/accp.h"
#include "accp_error.h"
#include "accp_internal.h"
#include "accp_kernels.h"

namespace {

constexpr int64_t NMAX = ACCP_NOALLOC_MAX_PARTICLES;
constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t IDX_BYTES = round256((size_t)NMAX * sizeof(int32_t));
constexpr size_t SMALL_BYTES = 256;  // counts, tile count, offsets: each its own 256-byte slot
constexpr size_t TILE_BYTES = round256((size_t)((NMAX + 63) / 64) * sizeof(accp::TileDesc));
constexpr size_t SCRATCH_BYTES = 2 * IDX_BYTES + 2 * SMALL_BYTES + TILE_BYTES + 2 * SMALL_BYTES;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};  // device address of scratch, per device

}  // namespace

extern "C" int accp_track_device_noalloc(accp_lattice *lat, double *d_r, int n, int nturn, int st, int ed,
                                         const accp_track_opts *opts, void *stream) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return accp::set_error(ACCP_E_HIP, "accp_track_device_noalloc: no current HIP device");
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return accp::set_error(ACCP_E_HIP, "accp_track_device_noalloc: hipGetSymbolAddress of the workspace failed");
      g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    base = g_scratch[dev];
  }
  accp::FixedWorkspace ws;
  unsigned char *q = base;
  ws.idx0 = reinterpret_cast<int32_t *>(q);
  q += IDX_BYTES;
  ws.idx1 = reinterpret_cast<int32_t *>(q);
  q += IDX_BYTES;
  ws.cnt0 = reinterpret_cast<int32_t *>(q);
  q += SMALL_BYTES;
  ws.cnt1 = reinterpret_cast<int32_t *>(q);
  q += SMALL_BYTES;
  ws.tiles = q;
  q += TILE_BYTES;
  ws.ntiles = reinterpret_cast<int32_t *>(q);
  q += SMALL_BYTES;
  ws.offs = reinterpret_cast<int64_t *>(q);
  ws.max_n = NMAX;
  return accp::track_device_fixed(lat, d_r, n, nturn, st, ed, opts, stream, ws);
}
