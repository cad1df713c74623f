// accp_wake.hip -- binned longitudinal resonator wake on gfx950.
//
// Per pass: (1) integer histogram of z over [zmin, zmax) -- each workgroup
// bins into an LDS histogram with LDS atomics, then adds its non-empty bins
// to the global counts (integer adds: exact and order-independent);
// (2) potential V_k = q_m * sum_{m>=k} cnt_m * W((m-k)*dz/c), one lane per
// bin; (3) delta -= V_bin / E0 for every live particle inside the window.
// Up to WAKE_FUSE_BINS bins, (2) and (3) are one kernel: every workgroup
// forms the whole potential in LDS from the counts, then kicks its share of
// the beam.  The kernel that reads the counts also clears them for the next
// pass (the last workgroup to finish reading), so a pass leaves no state
// behind and replays of a captured pass need no clearing launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "accp_table.h"
#include "accp_wake.h"

namespace accp {

static thread_local std::string g_wake_err;
const char *wake_last_error() { return g_wake_err.c_str(); }

double wake_function(double rs, double q, double fr, double tau) {
  const double wr = TWOPI * fr;
  const double alpha = wr / (2 * q);
  if (tau == 0) return alpha * rs;
  const double wb = std::sqrt(wr * wr - alpha * alpha);
  return 2 * alpha * rs * std::exp(-alpha * tau) * (std::cos(wb * tau) - alpha / wb * std::sin(wb * tau));
}

constexpr int WB = 256;
// bins up to which the potential is formed inside the kick kernel (each
// workgroup sums nb^2/2 terms)
constexpr int WAKE_FUSE_BINS = 512;
// particles per workgroup of the histogram and kick passes
constexpr int WAKE_PER_WG = 1024;

// The potential of bin k from the counts lc and the wake table lw[0..nb4)
// (zero from nb on, nb4 = nb rounded up to 4): the lags d in increasing
// order over four interleaved partial sums, every lane of the workgroup at
// the same lag (wave-uniform control flow, the table read is a broadcast).
// PADDED: lc holds 2 nb + 4 counts, zero from nb on; otherwise nb, and a
// bin past the window reads the last one and counts zero.
template <bool PADDED>
__device__ __forceinline__ double wake_potential(const uint32_t *lc, const double *lw, int nb, int nb4, int k) {
  auto count = [&](int m) -> double {
    if constexpr (PADDED) return (double)lc[m];
    else return m < nb ? (double)lc[m < nb ? m : nb - 1] : 0.0;
  };
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int d = 0; d < nb4; d += 4) {
    s0 += count(k + d) * lw[d];
    s1 += count(k + d + 1) * lw[d + 1];
    s2 += count(k + d + 2) * lw[d + 2];
    s3 += count(k + d + 3) * lw[d + 3];
  }
  return (s0 + s1) + (s2 + s3);
}

// Stages the table (zero-padded to nb4) and the counts (PADDED: zero-padded
// to 2 nb + 4) into LDS.
template <bool PADDED>
__device__ __forceinline__ void wake_stage(const uint32_t *cnt, const double *wt, int nb, int nb4, uint32_t *lc,
                                           double *lw) {
  for (int i = threadIdx.x; i < nb4; i += WB) lw[i] = i < nb ? wt[i] : 0.0;
  for (int i = threadIdx.x; i < (PADDED ? 2 * nb + 4 : nb); i += WB) lc[i] = i < nb ? cnt[i] : 0u;
}

// Called by every workgroup after it has read all nb counts (into LDS): the
// last one to get here clears the counts and the finish counter `done`.
__device__ __forceinline__ void clear_after_read(uint32_t *cnt, int nb, uint32_t *done) {
  __shared__ int last;
  __syncthreads();  // every lane's reads of cnt have returned
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last) {
    for (int i = threadIdx.x; i < nb; i += WB) cnt[i] = 0u;
    if (threadIdx.x == 0) atomicExch(done, 0u);
  }
}

__device__ __forceinline__ int wake_bin(const double *p, double zmin, double dz, int nb) {
  if (p[0] != p[0]) return -1;  // lost
  const double f = floor((p[4] - zmin) / dz);
  return (f >= 0 && f < nb) ? (int)f : -1;
}

__global__ __launch_bounds__(WB) void k_wake_hist(const double *__restrict__ r, int64_t n,
                                                  double zmin, double dz, int nb,
                                                  uint32_t *__restrict__ cnt) {
  extern __shared__ uint32_t h[];
  for (int i = threadIdx.x; i < nb; i += WB) h[i] = 0u;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * WB + threadIdx.x; i < n; i += (int64_t)gridDim.x * WB) {
    const int b = wake_bin(r + 6 * i, zmin, dz, nb);
    if (b >= 0) atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += WB)
    if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// LDS bytes of the potential's staging (table, counts, and the potential
// itself when the kick kernel forms it).
static size_t wake_lds_bytes(int nb, bool fused) {
  const int nb4 = (nb + 3) & ~3;
  return fused ? sizeof(double) * (nb4 + nb) + sizeof(uint32_t) * (2 * nb + 4)
               : sizeof(double) * nb4 + sizeof(uint32_t) * nb;
}

// V_k = qm * sum_{m >= k} cnt_m W((m - k) dz / c) (wake_potential); counts
// and the wake table are staged in LDS first (cnt holds nb counts and the
// finish counter after them).
__global__ __launch_bounds__(WB) void k_wake_conv(uint32_t *__restrict__ cnt,
                                                  const double *__restrict__ wt, int nb, double qm,
                                                  double *__restrict__ v) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int nb4 = (nb + 3) & ~3;
  double *lw = reinterpret_cast<double *>(lds_raw);
  uint32_t *lc = reinterpret_cast<uint32_t *>(lw + nb4);
  wake_stage<false>(cnt, wt, nb, nb4, lc, lw);
  clear_after_read(cnt, nb, cnt + nb);
  for (int k = blockIdx.x * WB + threadIdx.x; k < nb; k += gridDim.x * WB)
    v[k] = qm * wake_potential<false>(lc, lw, nb, nb4, k);
}

// The potential (as k_wake_conv, every workgroup the whole of it, in LDS)
// and the kick of particles [blockIdx.x * per_wg, +per_wg).
__global__ __launch_bounds__(WB) void k_wake_conv_kick(double *__restrict__ r, int64_t n, double zmin,
                                                       double dz, int nb, uint32_t *__restrict__ cnt,
                                                       const double *__restrict__ wt, double qm, double e0ev,
                                                       int64_t per_wg) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int nb4 = (nb + 3) & ~3;
  double *lw = reinterpret_cast<double *>(lds_raw);
  double *lv = lw + nb4;
  uint32_t *lc = reinterpret_cast<uint32_t *>(lv + nb);
  wake_stage<true>(cnt, wt, nb, nb4, lc, lw);
  clear_after_read(cnt, nb, cnt + nb);
  for (int k = threadIdx.x; k < nb; k += WB) lv[k] = qm * wake_potential<true>(lc, lw, nb, nb4, k);
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * per_wg, i1 = i0 + per_wg < n ? i0 + per_wg : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += WB) {
    double *p = r + 6 * i;
    const int b = wake_bin(p, zmin, dz, nb);
    if (b >= 0) p[5] -= lv[b] / e0ev;
  }
}

__global__ __launch_bounds__(WB) void k_wake_kick(double *__restrict__ r, int64_t n, double zmin,
                                                  double dz, int nb, const double *__restrict__ v,
                                                  double e0ev) {
  for (int64_t i = (int64_t)blockIdx.x * WB + threadIdx.x; i < n; i += (int64_t)gridDim.x * WB) {
    double *p = r + 6 * i;
    const int b = wake_bin(p, zmin, dz, nb);
    if (b >= 0) p[5] -= v[b] / e0ev;
  }
}

int wake_prepare(WakeState &w, const accp_elem *E, int n, const std::vector<int> &wakes) {
  wake_release(w);
  for (int idx : wakes) {
    if (idx < 0 || idx >= n) continue;
    const accp_elem &e = E[idx];
    if (e.type == ACCP_TRANSWAKE) {
      // ATransWake::pass constants (accel.cpp:1605-1615)
      WakeElem we;
      we.eidx = idx;
      we.kind = 1;
      we.len = e.len;
      we.qm_charge = e.wake_charge;
      we.e0ev = e.energy * 1e6;
      we.wx = e.tw_w[0];
      we.wy = e.tw_w[1];
      const double kRx = 2 * PI * e.tw_fr[0] * 1.0e9 / C0, kRy = 2 * PI * e.tw_fr[1] * 1.0e9 / C0;
      we.ax = kRx / 2 / e.tw_q[0];
      we.ay = kRy / 2 / e.tw_q[1];
      we.kbx = kRx * std::sqrt(1.0 - 0.25 / (e.tw_q[0] * e.tw_q[0]));
      we.kby = kRy * std::sqrt(1.0 - 0.25 / (e.tw_q[1] * e.tw_q[1]));
      w.elems.push_back(we);
      continue;
    }
    WakeElem we;
    we.eidx = idx;
    we.nbins = e.nbins;
    we.zmin = e.wake_zmin;
    we.dz = (e.wake_zmax - e.wake_zmin) / e.nbins;
    we.qm_charge = e.wake_charge;
    we.e0ev = e.energy * 1e6;
    std::vector<double> wt(e.nbins);
    for (int d = 0; d < e.nbins; ++d) wt[d] = wake_function(e.wake_rs, e.wake_q, e.wake_fr, d * we.dz / C0);
    if (hipMalloc(&we.d_wt, sizeof(double) * e.nbins) != hipSuccess ||
        hipMalloc(&we.d_cnt, sizeof(uint32_t) * (e.nbins + 1)) != hipSuccess ||
        hipMalloc(&we.d_v, sizeof(double) * e.nbins) != hipSuccess) {
      w.elems.push_back(we);
      wake_release(w);
      g_wake_err = "wake buffers: hipMalloc failed";
      return ACCP_E_NOMEM;
    }
    // the counts and the finish counter start cleared; every pass leaves them so
    if (hipMemcpy(we.d_wt, wt.data(), sizeof(double) * e.nbins, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(we.d_cnt, 0, sizeof(uint32_t) * (e.nbins + 1)) != hipSuccess) {
      w.elems.push_back(we);
      wake_release(w);
      g_wake_err = "wake table upload failed";
      return ACCP_E_HIP;
    }
    w.elems.push_back(we);
  }
  return ACCP_OK;
}

int wake_apply(WakeState &w, int eidx, double *d_r, int64_t n, hipStream_t s) {
  WakeElem *we = nullptr;
  for (WakeElem &x : w.elems)
    if (x.eidx == eidx) we = &x;
  if (!we) {
    g_wake_err = "no prepared wake at this element";
    return ACCP_E_STATE;
  }
  if (we->kind == 1) {
    const char *err = nullptr;
    const int rc = transwake_apply(*we, d_r, n, s, &err);
    if (rc) g_wake_err = err ? err : "TransWake failed";
    return rc;
  }
  if (n <= 0) return ACCP_OK;
  const int nb = we->nbins;
  const int grid = (int)std::min<int64_t>((n + WAKE_PER_WG - 1) / WAKE_PER_WG, 2048);
  hipLaunchKernelGGL(k_wake_hist, dim3(grid), dim3(WB), sizeof(uint32_t) * nb, s, d_r, n, we->zmin,
                     we->dz, nb, we->d_cnt);
  const double qm = we->qm_charge / (double)n;
  if (nb <= WAKE_FUSE_BINS) {
    const int64_t per_wg = (n + grid - 1) / grid;
    hipLaunchKernelGGL(k_wake_conv_kick, dim3(grid), dim3(WB), wake_lds_bytes(nb, true), s,
                       d_r, n, we->zmin, we->dz, nb, we->d_cnt, we->d_wt, qm, we->e0ev, per_wg);
  } else {
    hipLaunchKernelGGL(k_wake_conv, dim3((nb + WB - 1) / WB), dim3(WB), wake_lds_bytes(nb, false), s,
                       we->d_cnt, we->d_wt, nb, qm, we->d_v);
    hipLaunchKernelGGL(k_wake_kick, dim3(grid), dim3(WB), 0, s, d_r, n, we->zmin, we->dz, nb, we->d_v,
                       we->e0ev);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_wake_err = hipGetErrorString(e);
    return ACCP_E_HIP;
  }
  return ACCP_OK;
}

void wake_release(WakeState &w) {
  for (WakeElem &x : w.elems) {
    transwake_release(x);
    if (x.d_wt) (void)hipFree(x.d_wt);
    if (x.d_cnt) (void)hipFree(x.d_cnt);
    if (x.d_v) (void)hipFree(x.d_v);
  }
  w.elems.clear();
}

}  // namespace accp
