// accp_wake.hip -- binned longitudinal resonator wake on gfx950.
//
// Per pass: (1) integer histogram of z over [zmin, zmax) -- each workgroup
// bins into an LDS histogram with LDS atomics, then adds its non-empty bins
// to the global counts (integer adds: exact and order-independent);
// (2) potential V_k = q_m * sum_{m>=k} cnt_m * W((m-k)*dz/c), one lane per
// bin; (3) delta -= V_bin / E0 for every live particle inside the window.
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

__global__ __launch_bounds__(WB) void k_wake_zero(uint32_t *cnt, int nb) {
  for (int i = blockIdx.x * WB + threadIdx.x; i < nb; i += gridDim.x * WB) cnt[i] = 0u;
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

// V_k = qm * sum_{m >= k} cnt_m W((m - k) dz / c), summed in increasing m
// like the oracle; counts and the wake table are staged in LDS first.
__global__ __launch_bounds__(WB) void k_wake_conv(const uint32_t *__restrict__ cnt,
                                                  const double *__restrict__ wt, int nb, double qm,
                                                  double *__restrict__ v) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *lw = reinterpret_cast<double *>(lds_raw);
  uint32_t *lc = reinterpret_cast<uint32_t *>(lw + nb);
  for (int i = threadIdx.x; i < nb; i += WB) {
    lw[i] = wt[i];
    lc[i] = cnt[i];
  }
  __syncthreads();
  for (int k = blockIdx.x * WB + threadIdx.x; k < nb; k += gridDim.x * WB) {
    double s = 0;
    for (int m = k; m < nb; ++m) s += (double)lc[m] * lw[m - k];
    v[k] = qm * s;
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
        hipMalloc(&we.d_cnt, sizeof(uint32_t) * e.nbins) != hipSuccess ||
        hipMalloc(&we.d_v, sizeof(double) * e.nbins) != hipSuccess) {
      w.elems.push_back(we);
      wake_release(w);
      g_wake_err = "wake buffers: hipMalloc failed";
      return ACCP_E_NOMEM;
    }
    if (hipMemcpy(we.d_wt, wt.data(), sizeof(double) * e.nbins, hipMemcpyHostToDevice) != hipSuccess) {
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
  int grid = (int)std::min<int64_t>((n + WB - 1) / WB, 2048);
  hipLaunchKernelGGL(k_wake_zero, dim3((nb + WB - 1) / WB), dim3(WB), 0, s, we->d_cnt, nb);
  hipLaunchKernelGGL(k_wake_hist, dim3(grid), dim3(WB), sizeof(uint32_t) * nb, s, d_r, n, we->zmin,
                     we->dz, nb, we->d_cnt);
  hipLaunchKernelGGL(k_wake_conv, dim3((nb + WB - 1) / WB), dim3(WB), (sizeof(double) + sizeof(uint32_t)) * nb, s, we->d_cnt,
                     we->d_wt, nb,
                     we->qm_charge / (double)n, we->d_v);
  hipLaunchKernelGGL(k_wake_kick, dim3(grid), dim3(WB), 0, s, d_r, n, we->zmin, we->dz, nb, we->d_v,
                     we->e0ev);
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
