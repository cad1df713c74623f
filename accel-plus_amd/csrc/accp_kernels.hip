// accp_kernels.hip -- gfx950 kernels of the tracking backend.
//
// k_track: the persistent FP64 tracker (accp_track.h) walking the element
// table.  The same kernel serves the plain call (one seed, tiles implied by
// blockIdx) and the batched call (a tile list over many seeds).  Beside it:
// calc_beam_stat on monitor snapshots and the working-set bookkeeping
// between chunks of turns.
//
// Arithmetic is FP64 throughout; there is no dense contraction, so no MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "accp_kernels.h"
#include "accp_track.h"

namespace accp {

template <int BLOCK, int PPL, bool LDS, bool FULL>
__global__ __launch_bounds__(BLOCK) void k_track(KernelCtx ctx) {
  track_tile<BLOCK, PPL, LDS, FULL, TableWalker>(ctx);
}

// ---- calc_beam_stat on snapshots (utility.cpp:322-373) ----------------------
// One workgroup per record; two passes (mean, then central moments) with a
// fixed-order LDS tree reduction, so the result does not depend on timing.
constexpr int STAT_BLOCK = 256;

__device__ void block_sum(double (*buf)[STAT_BLOCK], int nv) {
  for (int s = STAT_BLOCK / 2; s > 0; s >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < s)
      for (int v = 0; v < nv; ++v) buf[v][threadIdx.x] += buf[v][threadIdx.x + s];
  }
  __syncthreads();
}

__global__ __launch_bounds__(STAT_BLOCK) void k_beam_stat(const double *__restrict__ snap,
                                                          int64_t n, double turn0, int interval,
                                                          double *__restrict__ out) {
  __shared__ double buf[13][STAT_BLOCK];
  const double *r = snap + (int64_t)blockIdx.x * n * 6;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t j = threadIdx.x; j < n; j += STAT_BLOCK) {
    const double *q = r + 6 * j;
    if (q[0] != q[0]) continue;
    s[6] += 1.0;
    for (int i = 0; i < 6; ++i) s[i] += q[i];
  }
  for (int v = 0; v < 7; ++v) buf[v][threadIdx.x] = s[v];
  block_sum(buf, 7);
  const double nl = buf[6][0];
  double mean[6];
  for (int i = 0; i < 6; ++i) mean[i] = buf[i][0] / nl;
  __syncthreads();
  const int pa[6] = {0, 2, 0, 1, 2, 3}, pb[6] = {1, 3, 5, 5, 4, 4};
  double q2[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t j = threadIdx.x; j < n; j += STAT_BLOCK) {
    const double *q = r + 6 * j;
    if (q[0] != q[0]) continue;
    double dv[6];
    for (int i = 0; i < 6; ++i) {
      dv[i] = q[i] - mean[i];
      q2[i] += dv[i] * dv[i];
    }
    for (int i = 0; i < 6; ++i) q2[6 + i] += dv[pa[i]] * dv[pb[i]];
  }
  for (int v = 0; v < 12; ++v) buf[v][threadIdx.x] = q2[v];
  block_sum(buf, 12);
  if (threadIdx.x == 0) {
    double *o = out + (int64_t)blockIdx.x * 20;
    o[0] = turn0 + (double)blockIdx.x * interval;
    o[1] = nl;
    for (int i = 0; i < 6; ++i) {
      o[2 + i] = mean[i];
      o[8 + i] = sqrt(buf[i][0] / nl);
      o[14 + i] = buf[6 + i][0] / nl;
    }
  }
}

// ---- working-set bookkeeping ---------------------------------------------------
__global__ __launch_bounds__(256) void k_chunk_init(int32_t *idx, int64_t n, int32_t *loss_turn,
                                                    int32_t nturn, int32_t *cnt,
                                                    const int64_t *offs, int nseed) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    idx[i] = (int32_t)i;
    if (loss_turn) loss_turn[i] = nturn;
  }
  if (blockIdx.x == 0)
    for (int s = threadIdx.x; s < nseed; s += 256) cnt[s] = (int32_t)(offs[s + 1] - offs[s]);
}

// Tile list from per-seed working-set sizes: seed s owns positions
// [offs[s], offs[s] + cnt[s]) of the index list.
__global__ __launch_bounds__(256) void k_make_tiles(const int32_t *cnt, const int64_t *offs,
                                                    int nseed, int per, TileDesc *tiles,
                                                    int32_t *ntiles) {
  // Seeds in chunks of 256: an inclusive scan of their tile counts, then
  // every thread writes tiles of the chunk (tile t belongs to the first seed
  // whose inclusive count exceeds t).  Seed-major, tile order fixed.
  __shared__ int32_t base[256];
  __shared__ int32_t cnts[256];
  __shared__ int64_t offs_s[256];
  int32_t total = 0;
  for (int s0 = 0; s0 < nseed; s0 += 256) {
    const int s = s0 + threadIdx.x;
    const int32_t c = s < nseed ? cnt[s] : 0;
    cnts[threadIdx.x] = c;
    offs_s[threadIdx.x] = s < nseed ? offs[s] : 0;
    base[threadIdx.x] = (c + per - 1) / per;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int32_t v = threadIdx.x >= (unsigned)d ? base[threadIdx.x - d] : 0;
      __syncthreads();
      base[threadIdx.x] += v;
      __syncthreads();
    }
    const int32_t chunk = base[255];
    for (int32_t t = threadIdx.x; t < chunk; t += 256) {
      int lo = 0, hi = 255;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] > t) hi = mid;
        else lo = mid + 1;
      }
      const int32_t k = t - (lo ? base[lo - 1] : 0);
      const int32_t left = cnts[lo] - k * per;
      tiles[total + t] = TileDesc{s0 + lo, left < per ? left : per, offs_s[lo] + (int64_t)k * per};
    }
    total += chunk;
    __syncthreads();
  }
  if (threadIdx.x == 0) *ntiles = total;
}

__global__ __launch_bounds__(256) void k_zero_i32(int32_t *p, int n) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// Keep the live particles of each tile: one atomic per wave on the seed's
// counter, lanes placed by ballot prefix.  Which slot a particle lands in
// depends on timing, its trajectory does not.
__global__ __launch_bounds__(256) void k_compact(const TileDesc *tiles, const int32_t *ntiles,
                                                 const int32_t *idx_in, const double *r,
                                                 const int64_t *offs, int32_t *cnt_out,
                                                 int32_t *idx_out) {
  if ((int)blockIdx.x >= *ntiles) return;
  const TileDesc t = tiles[blockIdx.x];
  const int lane = __lane_id();
  for (int j0 = 0; j0 < t.count; j0 += 256) {
    const int j = j0 + threadIdx.x;
    int32_t i = -1;
    bool live = false;
    if (j < t.count) {
      i = idx_in[t.start + j];
      const double x = r[6 * (int64_t)i];
      live = x == x;
    }
    const unsigned long long m = __ballot(live);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&cnt_out[t.lat], __popcll(m));
    base = __shfl(base, 0);
    if (live) idx_out[offs[t.lat] + base + __popcll(m & ((1ull << lane) - 1ull))] = i;
  }
}

// ---- launchers -------------------------------------------------------------------
template <int BLOCK, int PPL, bool LDS>
static hipError_t launch_one(const KernelCtx &ctx, int grid, size_t lds, bool full, hipStream_t s) {
  if (full)
    hipLaunchKernelGGL((k_track<BLOCK, PPL, LDS, true>), dim3(grid), dim3(BLOCK), lds, s, ctx);
  else
    hipLaunchKernelGGL((k_track<BLOCK, PPL, LDS, false>), dim3(grid), dim3(BLOCK), lds, s, ctx);
  return hipGetLastError();
}

hipError_t launch_track(const KernelCtx &ctx, int grid, int block, int ppl, bool lds,
                        size_t lds_bytes, bool full, hipStream_t s) {
#define ACCP_CASE(B, P)                                                        \
  if (block == B && ppl == P)                                                  \
    return lds ? launch_one<B, P, true>(ctx, grid, lds_bytes, full, s)        \
               : launch_one<B, P, false>(ctx, grid, 0, full, s);
  ACCP_CASE(64, 1)
  ACCP_CASE(128, 1)
  ACCP_CASE(256, 1)
  ACCP_CASE(64, 2)
  ACCP_CASE(128, 2)
  ACCP_CASE(256, 2)
#undef ACCP_CASE
  return hipErrorInvalidValue;
}

__global__ void k_set_offsets(int64_t *offs, int64_t n) {
  if (threadIdx.x == 0) {
    offs[0] = 0;
    offs[1] = n;
  }
}

hipError_t launch_set_offsets(int64_t *offs, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_set_offsets, dim3(1), dim3(64), 0, s, offs, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_loss_init(const double *r, int64_t n, int32_t *loss_turn, int32_t nturn) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    loss_turn[i] = r[6 * i] == r[6 * i] ? nturn : -1;
}

hipError_t launch_loss_init(const double *r, int64_t n, int32_t *loss_turn, int32_t nturn, hipStream_t s) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
  if (grid > 0) hipLaunchKernelGGL(k_loss_init, dim3(grid), dim3(256), 0, s, r, n, loss_turn, nturn);
  return hipGetLastError();
}

hipError_t launch_chunk_init(int32_t *idx, int64_t n, int32_t *loss_turn, int32_t nturn,
                             int32_t *cnt, const int64_t *offs, int nseed, hipStream_t s) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256 + 1, 4096);
  hipLaunchKernelGGL(k_chunk_init, dim3(grid), dim3(256), 0, s, idx, n, loss_turn, nturn, cnt, offs, nseed);
  return hipGetLastError();
}

hipError_t launch_make_tiles(const int32_t *cnt, const int64_t *offs, int nseed, int per,
                             TileDesc *tiles, int32_t *ntiles, hipStream_t s) {
  hipLaunchKernelGGL(k_make_tiles, dim3(1), dim3(256), 0, s, cnt, offs, nseed, per, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_compact(const TileDesc *tiles, const int32_t *ntiles, int max_tiles,
                          const int32_t *idx_in, const double *r, const int64_t *offs,
                          int32_t *cnt_out, int32_t *idx_out, int nseed, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_i32, dim3((nseed + 255) / 256), dim3(256), 0, s, cnt_out, nseed);
  if (max_tiles > 0)
    hipLaunchKernelGGL(k_compact, dim3(max_tiles), dim3(256), 0, s, tiles, ntiles, idx_in, r, offs,
                       cnt_out, idx_out);
  return hipGetLastError();
}

hipError_t launch_beam_stat(const double *snap, int64_t n, int nrec, double turn0, int interval,
                            double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_beam_stat, dim3(nrec), dim3(STAT_BLOCK), 0, s, snap, n, turn0, interval,
                     out);
  return hipGetLastError();
}

}  // namespace accp
