// accp_wake.h -- collective elements: the longitudinal resonator wake (no
// reference counterpart) and the transverse resonator wake of the whole beam
// (the reference's pairwise ATransWake, src/lib/accel/accel.cpp:1603-1657).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/accp.h"

namespace accp {

constexpr int WAKE_MAX_BINS = 4096;

struct WakeElem {
  int eidx = -1;
  int kind = 0;  // 0: LongWake (binned), 1: TransWake (sorted scan)
  int nbins = 0;
  double zmin = 0, dz = 0, qm_charge = 0, e0ev = 0;
  double *d_wt = nullptr;    // W(d*dz/c), d = 0..nbins-1 (W(0) = alpha*Rs)
  uint32_t *d_cnt = nullptr; // histogram of the current pass, then a finish counter (nbins + 1);
                             // cleared between passes by the kernel that reads them
  double *d_v = nullptr;     // wake potential per bin [V]
  // TransWake: length, strengths, decay and wave numbers per plane
  double len = 0, wx = 0, wy = 0, ax = 0, ay = 0, kbx = 0, kby = 0;
  void *tw_key[2] = {nullptr, nullptr}, *tw_idx[2] = {nullptr, nullptr}, *tw_term[2] = {nullptr, nullptr};
  void *tw_tmp = nullptr;
  size_t tw_cap[7] = {0, 0, 0, 0, 0, 0, 0};
};

struct WakeState {
  std::vector<WakeElem> elems;
};

// Resonator wake W(tau) [V/C]: 2*alpha*Rs*exp(-alpha*tau)*(cos(wb*tau) -
// alpha/wb*sin(wb*tau)) for tau > 0, alpha*Rs at tau = 0 (half the field a
// charge sees from itself), alpha = wr/(2Q), wb = sqrt(wr^2 - alpha^2).
double wake_function(double rs, double q, double fr, double tau);

int wake_prepare(WakeState &w, const accp_elem *E, int n, const std::vector<int> &wakes);
// Histogram, potential and energy kick of element `eidx` on the AoS beam.
int wake_apply(WakeState &w, int eidx, double *d_r, int64_t n, hipStream_t s);
void wake_release(WakeState &w);
// TransWake pass on the AoS beam (accp_twake.hip).
int transwake_apply(WakeElem &w, double *d_r, int64_t n, hipStream_t s, const char **err);
void transwake_release(WakeElem &w);
const char *wake_last_error();

}  // namespace accp
