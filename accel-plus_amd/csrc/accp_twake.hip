// accp_twake.hip -- transverse resonator wake of the whole beam on gfx950.
//
// ATransWake (src/lib/accel/accel.cpp:1603-1657) kicks particle i by
//   dx'_i = -q W sum_{j: z_j >= z_i, j != i} exp(a (z_i - z_j)) sin(kb (z_i - z_j)) x_j / E
// (and the same in y), a pairwise O(N^2) sum over the OpenMP slice a thread
// holds.  exp(a d) sin(kb d) is the imaginary part of exp(c d), c = a + i kb,
// so with the live particles sorted by z (leaders first) the sum for the
// p-th particle is Im T_p of the linear recurrence
//   T_p = T_{p-1} exp(c (z_p - z_{p-1})) + x_p,
// whose factors have modulus <= 1.  The recurrence is an associative scan
// of (multiplier, value) pairs: one stable radix sort and one deterministic
// scan per pass, O(N log N), the same sum up to round-off, over the whole
// beam (the reference's per-slice coupling is not reproduced).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cfloat>
#include <cmath>

#include "accp_table.h"
#include "accp_wake.h"

namespace accp {

namespace {

constexpr int TB = 256;

// x and y planes: multiplier m and value v, complex
struct TwTerm {
  double mxr, mxi, vxr, vxi, myr, myi, vyr, vyi;
};

struct TwCombine {
  // a then b: T -> (T m_a + v_a) m_b + v_b
  __device__ __host__ TwTerm operator()(const TwTerm &a, const TwTerm &b) const {
    TwTerm r;
    r.mxr = a.mxr * b.mxr - a.mxi * b.mxi;
    r.mxi = a.mxr * b.mxi + a.mxi * b.mxr;
    r.vxr = a.vxr * b.mxr - a.vxi * b.mxi + b.vxr;
    r.vxi = a.vxr * b.mxi + a.vxi * b.mxr + b.vxi;
    r.myr = a.myr * b.myr - a.myi * b.myi;
    r.myi = a.myr * b.myi + a.myi * b.myr;
    r.vyr = a.vyr * b.myr - a.vyi * b.myi + b.vyr;
    r.vyi = a.vyr * b.myi + a.vyi * b.myr + b.vyi;
    return r;
  }
};

__global__ __launch_bounds__(TB) void k_tw_keys(const double *__restrict__ r, int64_t n, double *__restrict__ key,
                                               int32_t *__restrict__ idx) {
  for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) {
    const double *p = r + 6 * i;
    key[i] = p[0] != p[0] ? -DBL_MAX : p[4];  // lost particles sort last
    idx[i] = (int32_t)i;
  }
}

__global__ __launch_bounds__(TB) void k_tw_terms(const double *__restrict__ r, int64_t n,
                                                const double *__restrict__ key, const int32_t *__restrict__ idx,
                                                double ax, double kbx, double ay, double kby, TwTerm *__restrict__ t) {
  for (int64_t q = (int64_t)blockIdx.x * TB + threadIdx.x; q < n; q += (int64_t)gridDim.x * TB) {
    const double *p = r + 6 * idx[q];
    TwTerm u{1, 0, 0, 0, 1, 0, 0, 0};
    if (p[0] == p[0]) {
      u.vxr = p[0];
      u.vyr = p[2];
      if (q == 0) {
        u.mxr = u.myr = 0;
      } else {
        const double d = key[q] - key[q - 1];  // <= 0: the predecessor leads
        double s, c;
        sincos(kbx * d, &s, &c);
        const double ex = exp(ax * d);
        u.mxr = ex * c;
        u.mxi = ex * s;
        sincos(kby * d, &s, &c);
        const double ey = exp(ay * d);
        u.myr = ey * c;
        u.myi = ey * s;
      }
    }
    t[q] = u;
  }
}

__global__ __launch_bounds__(TB) void k_tw_kick(double *__restrict__ r, int64_t n, const int32_t *__restrict__ idx,
                                               const TwTerm *__restrict__ t, double cx, double cy) {
  for (int64_t q = (int64_t)blockIdx.x * TB + threadIdx.x; q < n; q += (int64_t)gridDim.x * TB) {
    double *p = r + 6 * idx[q];
    if (p[0] != p[0]) continue;
    p[1] += -cx * t[q].vxi * (1 + p[5]);
    p[3] += -cy * t[q].vyi * (1 + p[5]);
  }
}

__global__ __launch_bounds__(TB) void k_tw_drift(double *__restrict__ r, int64_t n, double half) {
  for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) {
    double *p = r + 6 * i;
    if (p[0] != p[0]) continue;
    const double nl = half / (1 + p[5]);  // fastdrift (atpass.cpp:416-428)
    p[0] += nl * p[1];
    p[2] += nl * p[3];
    p[4] -= nl * (p[1] * p[1] + p[3] * p[3]) / (2 * (1 + p[5]));
  }
}

int grow(void *&ptr, size_t &cap, size_t bytes) {
  if (bytes <= cap) return 0;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  if (hipMalloc(&ptr, bytes) != hipSuccess) return -1;
  cap = bytes;
  return 0;
}

}  // namespace

int transwake_apply(WakeElem &w, double *d_r, int64_t n, hipStream_t s, const char **err) {
  if (n <= 0) return ACCP_OK;
  const int grid = (int)std::min<int64_t>((n + TB - 1) / TB, 4096);
  const size_t nk = (size_t)n;
  if (grow(w.tw_key[0], w.tw_cap[0], nk * sizeof(double)) || grow(w.tw_key[1], w.tw_cap[1], nk * sizeof(double)) ||
      grow(w.tw_idx[0], w.tw_cap[2], nk * sizeof(int32_t)) || grow(w.tw_idx[1], w.tw_cap[3], nk * sizeof(int32_t)) ||
      grow(w.tw_term[0], w.tw_cap[4], nk * sizeof(TwTerm)) || grow(w.tw_term[1], w.tw_cap[5], nk * sizeof(TwTerm))) {
    *err = "TransWake buffers: hipMalloc failed";
    return ACCP_E_NOMEM;
  }
  double *k0 = (double *)w.tw_key[0], *k1 = (double *)w.tw_key[1];
  int32_t *i0 = (int32_t *)w.tw_idx[0], *i1 = (int32_t *)w.tw_idx[1];
  TwTerm *t0 = (TwTerm *)w.tw_term[0], *t1 = (TwTerm *)w.tw_term[1];
  size_t sort_bytes = 0, scan_bytes = 0;
  if (rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, k0, k1, i0, i1, nk, 0, 64, s) != hipSuccess ||
      rocprim::deterministic_inclusive_scan(nullptr, scan_bytes, t0, t1, nk, TwCombine(), s) != hipSuccess ||
      grow(w.tw_tmp, w.tw_cap[6], std::max(sort_bytes, scan_bytes))) {
    *err = "TransWake: temporary storage";
    return ACCP_E_NOMEM;
  }
  if (w.len != 0) hipLaunchKernelGGL(k_tw_drift, dim3(grid), dim3(TB), 0, s, d_r, n, 0.5 * w.len);
  hipLaunchKernelGGL(k_tw_keys, dim3(grid), dim3(TB), 0, s, d_r, n, k0, i0);
  size_t sb = w.tw_cap[6];
  if (rocprim::radix_sort_pairs_desc(w.tw_tmp, sb, k0, k1, i0, i1, nk, 0, 64, s) != hipSuccess) {
    *err = "TransWake: sort failed";
    return ACCP_E_HIP;
  }
  hipLaunchKernelGGL(k_tw_terms, dim3(grid), dim3(TB), 0, s, d_r, n, k1, i1, w.ax, w.kbx, w.ay, w.kby, t0);
  sb = w.tw_cap[6];
  if (rocprim::deterministic_inclusive_scan(w.tw_tmp, sb, t0, t1, nk, TwCombine(), s) != hipSuccess) {
    *err = "TransWake: scan failed";
    return ACCP_E_HIP;
  }
  // q = Ne/N * 1000 pC per macro-particle (accel.cpp:1605); / (E * 1e6)
  const double q = w.qm_charge / (double)n * 1000;
  hipLaunchKernelGGL(k_tw_kick, dim3(grid), dim3(TB), 0, s, d_r, n, i1, t1, q * w.wx / w.e0ev,
                     q * w.wy / w.e0ev);
  if (w.len != 0) hipLaunchKernelGGL(k_tw_drift, dim3(grid), dim3(TB), 0, s, d_r, n, 0.5 * w.len);
  if (hipGetLastError() != hipSuccess) {
    *err = "TransWake: launch failed";
    return ACCP_E_HIP;
  }
  return ACCP_OK;
}

void transwake_release(WakeElem &w) {
  for (void *&p : w.tw_key) { if (p) (void)hipFree(p); p = nullptr; }
  for (void *&p : w.tw_idx) { if (p) (void)hipFree(p); p = nullptr; }
  for (void *&p : w.tw_term) { if (p) (void)hipFree(p); p = nullptr; }
  if (w.tw_tmp) (void)hipFree(w.tw_tmp);
  w.tw_tmp = nullptr;
  for (size_t &c : w.tw_cap) c = 0;
}

}  // namespace accp
