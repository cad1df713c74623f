// accp_naff.h -- NAFF (numerical analysis of fundamental frequencies) of many
// turn-by-turn signals on the GPU: one workgroup per signal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace accp {

// Signal k of a launch: samples sig[base + i*stride], i = 0..m-1, m a power
// of two in [NAFF_MIN_LEN, NAFF_MAX_LEN]; the search window [xlow, xhi] in
// tune units (naff, utility.cpp:48-96, clamps it to [0.9/m, 1/2]).
struct NaffJob {
  int64_t base;
  int64_t stride;
  int32_t m;
  int32_t pad;
  double xlow, xhi;
};

constexpr int NAFF_MIN_LEN = 4;
constexpr int NAFF_MAX_LEN = 4096;  // m doubles of LDS per workgroup (32 KB)

// Runs the jobs on stream s (synchronous: h_out holds the tunes on return).
// NaN samples give a deterministic result, as the reference's comparisons
// with NaN do (the spectrum search keeps its first bin, the section search
// always shrinks from above).
hipError_t naff_run(const double *d_sig, const NaffJob *h_jobs, int njobs, double *h_out, hipStream_t s);

}  // namespace accp
