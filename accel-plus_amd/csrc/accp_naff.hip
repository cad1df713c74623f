// accp_naff.hip -- NAFF of turn-by-turn signals on gfx950.
//
// naff (utility.cpp:48-96) finds the strongest line of a signal's spectrum
// inside a window and refines it by a golden-section search of the
// sin^4-windowed projection |sum x_i w_i exp(-2 pi i nu j)| over one bin on
// either side.  Here one 256-lane workgroup takes one signal:
//   1. the spectrum: the m real samples packed as m/2 complex points
//      (x_2j + i x_2j+1), a radix-2 FFT in LDS, and the textbook split into
//      the transform of the real signal;
//   2. the strongest bin of the window, first maximum as the reference's
//      sequential scan picks it (its pair 0 holds F_0 and F_m/2 together);
//   3. the golden-section search with the reference's ratio, placement and
//      stopping rule (mins.h:72-110), each projection a workgroup-wide sum
//      in a fixed order (deterministic from run to run).
// FP64 throughout; the projection sums dominate (about 25 evaluations of m
// sincos each).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "accp_naff.h"

namespace accp {

namespace {

constexpr int NB = 256;
constexpr int NW = NB / 64;
constexpr double NAFF_PI = 3.1415926535897932384626;  // utility.h:10
constexpr int GOLDEN_MAX_ITER = 300;  // the bracket shrinks below tolerance in < 60

// Sum of a and b over the workgroup: 64-lane butterflies, then the waves in a
// fixed order.  Every thread returns the same values.
__device__ void block_sum2(double &a, double &b, double *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = (red[0] + red[2]) + (red[4] + red[6]);
  b = (red[1] + red[3]) + (red[5] + red[7]);
}

// The larger value wins; equal values keep the lower bin.
__device__ void better(double &v, int &k, double v2, int k2) {
  if (v2 > v || (v2 == v && k2 < k)) {
    v = v2;
    k = k2;
  }
}

// -|projection| of the windowed signal xw (x_i w_i in LDS) on frequency nu
// (fprojection, utility.cpp:21-41).
__device__ double projection(const double *xw, int m, double nu, double *red) {
  double ry = 0, iy = 0;
  for (int i = threadIdx.x; i < m; i += NB) {
    double s, c;
    sincos(2 * NAFF_PI * nu * i, &s, &c);
    ry += c * xw[i];
    iy -= s * xw[i];
  }
  block_sum2(ry, iy, red);
  return -sqrt(ry * ry + iy * iy);
}

__global__ __launch_bounds__(NB) void k_naff(const double *__restrict__ sig, const NaffJob *__restrict__ jobs,
                                            double *__restrict__ out) {
  extern __shared__ __align__(16) double lds[];  // m doubles: m/2 complex points, then x_i w_i
  __shared__ double red[2 * NW];
  __shared__ double bestv[NW];
  __shared__ int bestk[NW];
  const NaffJob J = jobs[blockIdx.x];
  const int m = J.m, h = m >> 1;
  const int logh = 31 - __builtin_clz(h);
  const double *x = sig + J.base;
  const int64_t st = J.stride;
  double2 *z = reinterpret_cast<double2 *>(lds);

  // 1. packed complex points in bit-reversed order, then the butterflies
  for (int j = threadIdx.x; j < h; j += NB) {
    const int r = (int)(__brev((unsigned)j) >> (32 - logh));
    z[r] = make_double2(x[(2 * j) * st], x[(2 * j + 1) * st]);
  }
  __syncthreads();
  for (int len = 2; len <= h; len <<= 1) {
    const int half = len >> 1;
    for (int b = threadIdx.x; b < (h >> 1); b += NB) {
      const int k = b & (half - 1);
      const int p = (b / half) * len + k, q = p + half;
      double s, c;
      sincospi(-2.0 * k / len, &s, &c);
      const double2 u = z[p], v = z[q];
      const double tr = v.x * c - v.y * s, ti = v.x * s + v.y * c;
      z[p] = make_double2(u.x + tr, u.y + ti);
      z[q] = make_double2(u.x - tr, u.y - ti);
    }
    __syncthreads();
  }

  // 2. window and strongest bin (utility.cpp:54-72); bin k is the
  //    reference's pair 2k
  double xlow = J.xlow, xhi = J.xhi;
  xlow = xlow > 0.9 / m ? xlow : 0.9 / m;
  xhi = 0.5 < xhi ? 0.5 : xhi;
  // the host passes 0 <= xlow < xhi; bins stay inside [0, m/2)
  const int klo = min((((int)(xlow * m * 2) >> 1) << 1) >> 1, h - 1);
  const int khi = min((((int)(xhi * m * 2) >> 1) << 1) >> 1, h);
  auto power = [&](int k) -> double {
    if (k == 0) {  // pair 0: F_0 and F_m/2
      const double2 z0 = z[0];
      const double f0 = z0.x + z0.y, fh = z0.x - z0.y;
      return f0 * f0 + fh * fh;
    }
    const double2 a = z[k], b = z[h - k];
    // F_k = E_k + exp(-2 pi i k / m) O_k, E = (Z_k + conj Z_h-k)/2,
    // O = (Z_k - conj Z_h-k)/(2i)
    const double er = 0.5 * (a.x + b.x), ei = 0.5 * (a.y - b.y);
    const double orr = 0.5 * (a.y + b.y), oi = -0.5 * (a.x - b.x);
    double s, c;
    sincospi(-2.0 * k / m, &s, &c);
    const double fr = er + c * orr - s * oi, fi = ei + c * oi + s * orr;
    return fr * fr + fi * fi;
  };
  const double p0 = power(klo);
  int kbest = klo;
  if (p0 == p0) {  // a NaN first bin is kept, as the sequential scan keeps it
    double v = -DBL_MAX;
    int k = 0x7fffffff;
    for (int kk = klo + (int)threadIdx.x; kk < (khi > klo ? khi : klo + 1); kk += NB) better(v, k, power(kk), kk);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) better(v, k, __shfl_xor(v, off), __shfl_xor(k, off));
    if ((threadIdx.x & 63) == 0) {
      bestv[threadIdx.x >> 6] = v;
      bestk[threadIdx.x >> 6] = k;
    }
    __syncthreads();
    v = bestv[0];
    k = bestk[0];
    for (int w = 1; w < NW; ++w) better(v, k, bestv[w], bestk[w]);
    kbest = k;
  }
  __syncthreads();  // spectrum reads done before the buffer is reused

  // 3. windowed signal, bracket of one bin either side, golden section
  for (int i = threadIdx.x; i < m; i += NB) {
    const double w = sin(NAFF_PI * i / m);
    const double w2 = w * w;
    lds[i] = x[i * st] * (w2 * w2);
  }
  __syncthreads();
  const double imx = 2.0 * kbest;
  const double lo = (imx - 2) * 1.0 / m / 2.0, hi = (imx + 2) * 1.0 / m / 2.0;
  xlow = xlow > lo ? xlow : lo;
  xhi = xhi < hi ? xhi : hi;
  // Golden-section search for the minimum of -|projection| on [xlow, xhi]:
  // the bracket [a, d] holds two probes b < c; each step keeps the side of
  // the better probe, moves the bracket's far end onto the worse probe and
  // places one new probe a golden fraction into the longer part, until the
  // bracket is narrower than 3e-8 of |b| + |c|.  The ratio (0.61803399),
  // the first probes (the midpoint and a point into the longer half) and
  // the stopping rule are the reference's (Golden::minimize, mins.h:75-108,
  // called from naff, utility.cpp:74), so the tunes agree with its to the
  // stopping tolerance.
  const double g = 0.61803399, gc = 1.0 - g, tol = 3.0e-8;
  const double mid = (xlow + xhi) / 2.0;
  double a = xlow, d = xhi;
  const bool right_longer = fabs(d - mid) > fabs(mid - a);
  double b = right_longer ? mid : mid - gc * (mid - a);
  double c = right_longer ? mid + gc * (d - mid) : mid;
  double fb = projection(lds, m, b, red), fc = projection(lds, m, c, red);
  for (int it = 0; it < GOLDEN_MAX_ITER && fabs(d - a) > tol * (fabs(b) + fabs(c)); ++it) {
    if (fc < fb) {  // the minimum lies in [b, d]
      a = b;
      b = c;
      fb = fc;
      c = g * c + gc * d;
      fc = projection(lds, m, c, red);
    } else {  // the minimum lies in [a, c]
      d = c;
      c = b;
      fc = fb;
      b = g * b + gc * a;
      fb = projection(lds, m, b, red);
    }
  }
  if (threadIdx.x == 0) out[blockIdx.x] = fb < fc ? b : c;
}

}  // namespace

hipError_t naff_run(const double *d_sig, const NaffJob *h_jobs, int njobs, double *h_out, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  int mmax = 0;
  for (int k = 0; k < njobs; ++k) mmax = h_jobs[k].m > mmax ? h_jobs[k].m : mmax;
  NaffJob *d_jobs = nullptr;
  double *d_out = nullptr;
  hipError_t e = hipMalloc(&d_jobs, sizeof(NaffJob) * njobs);
  if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * njobs);
  if (e == hipSuccess) e = hipMemcpyAsync(d_jobs, h_jobs, sizeof(NaffJob) * njobs, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_naff, dim3(njobs), dim3(NB), sizeof(double) * mmax, s, d_sig, d_jobs, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h_out, d_out, sizeof(double) * njobs, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (d_jobs) (void)hipFree(d_jobs);
  if (d_out) (void)hipFree(d_out);
  return e;
}

}  // namespace accp
