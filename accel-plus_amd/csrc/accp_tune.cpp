// accp_tune.cpp -- tune diagnostics on the GPU: NAFF of turn-by-turn
// positions (tracktune, trackDiffusion, trackdnudexy of ALine).
//
// Every particle of a call is tracked in one launch with the state recorded
// at each turn start (accp::track_record); every signal (particle x plane x
// half) is then one workgroup of the NAFF kernel (accp_naff.hip).  The host
// keeps the reference's arithmetic around them (windows, diffusion rate,
// least-squares slopes), built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/accp.h"
#include "accp_error.h"
#include "accp_internal.h"
#include "accp_naff.h"

namespace {

constexpr double PI = 3.1415926535897932384626;  // accel.h:23

// Cap on the turn-by-turn record of one tracking launch (trackDiffusion
// tracks its particles in groups below it).
constexpr int64_t RECORD_BYTES = int64_t(8) << 30;

int fail(int code, const std::string &msg) { return accp::set_error(code, msg.c_str()); }

// The power of two that tracktune, trackDiffusion and trackdnudexy pad
// nturn to (aline.cpp:1169-1170).
unsigned pow2_at_least(unsigned n) {
  unsigned nt = 1;
  while (nt < n) nt <<= 1;
  return nt;
}

int check_window(double lo, double hi) {
  if (!(lo >= 0 && lo < hi && std::isfinite(hi))) return fail(ACCP_E_ARG, "tune window needs 0 <= low < high");
  return ACCP_OK;
}

int check_len(int m) {
  if (m < accp::NAFF_MIN_LEN || m > accp::NAFF_MAX_LEN || (m & (m - 1)))
    return fail(ACCP_E_ARG, "NAFF signals must have a power-of-two length from " +
                                std::to_string(accp::NAFF_MIN_LEN) + " to " + std::to_string(accp::NAFF_MAX_LEN));
  return ACCP_OK;
}

accp::NaffJob job(int64_t base, int64_t stride, int m, double lo, double hi) {
  accp::NaffJob j;
  j.base = base;
  j.stride = stride;
  j.m = m;
  j.pad = 0;
  j.xlow = lo;
  j.xhi = hi;
  return j;
}

int run_jobs(const double *d_sig, const std::vector<accp::NaffJob> &jobs, std::vector<double> &out) {
  out.assign(jobs.size(), 0.0);
  const hipError_t e = accp::naff_run(d_sig, jobs.data(), (int)jobs.size(), out.data(), nullptr);
  if (e != hipSuccess) return fail(ACCP_E_HIP, std::string("NAFF kernel: ") + hipGetErrorString(e));
  return ACCP_OK;
}

// Least-squares slope of y on x, centred (fitab.h:37-57).
double slope(int n, const double *x, const double *y) {
  double sx = 0, st2 = 0, b = 0;
  for (int i = 0; i < n; i++) sx += x[i];
  const double sxoss = sx / n;
  for (int i = 0; i < n; i++) {
    const double t = x[i] - sxoss;
    st2 += t * t;
    b += t * y[i];
  }
  return b / st2;
}

}  // namespace

extern "C" int accp_naff(const double *data, int nsig, int m, double xlow, double xhi, double *tunes) {
  if (!data || !tunes || nsig < 0) return fail(ACCP_E_ARG, "accp_naff: bad arguments");
  int rc;
  if ((rc = check_len(m)) || (rc = check_window(xlow, xhi))) return rc;
  if (nsig == 0) return ACCP_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(ACCP_E_HIP, "no HIP device available");
  const size_t bytes = sizeof(double) * (size_t)nsig * m;
  double *d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return fail(ACCP_E_NOMEM, "accp_naff: hipMalloc failed");
  std::vector<accp::NaffJob> jobs;
  for (int s = 0; s < nsig; ++s) jobs.push_back(job((int64_t)s * m, 1, m, xlow, xhi));
  std::vector<double> out;
  if (hipMemcpy(d, data, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(ACCP_E_HIP, "accp_naff: copy");
  if (!rc) rc = run_jobs(d, jobs, out);
  (void)hipFree(d);
  if (!rc)
    for (int s = 0; s < nsig; ++s) tunes[s] = out[s];
  return rc;
}

extern "C" int accp_tracktune(accp_lattice *L, int nturn, double x0, double y0, double xlow, double xhi,
                              double ylow, double yhi, double tune[2]) {
  if (!L || !tune || nturn < 1) return fail(ACCP_E_ARG, "accp_tracktune: bad arguments");
  const int nt = (int)pow2_at_least((unsigned)nturn);
  int rc;
  if ((rc = check_len(nt)) || (rc = check_window(xlow, xhi)) || (rc = check_window(ylow, yhi))) return rc;
  // aline.cpp:1159-1186: positions at the start and after each of nt - 1 turns
  double r[6] = {x0, 0, y0, 0, 0, 0};
  const double *rec = nullptr;
  if ((rc = accp::track_record(L, r, 1, nt, &rec))) return rc;
  std::vector<double> out;
  if ((rc = run_jobs(rec, {job(0, 6, nt, xlow, xhi), job(2, 6, nt, ylow, yhi)}, out))) return rc;
  tune[0] = out[0];
  tune[1] = out[1];
  return ACCP_OK;
}

extern "C" int accp_trackdiffusion(accp_lattice *L, int nturn, int np, const double *x0, const double *y0,
                                   double *nux, double *nuy, double *dif, double xlow, double xhi, double ylow,
                                   double yhi, double *total) {
  if (!L || np < 0 || (np > 0 && (!x0 || !y0 || !nux || !nuy || !dif)) || nturn < 1)
    return fail(ACCP_E_ARG, "accp_trackdiffusion: bad arguments");
  const int nt = (int)pow2_at_least((unsigned)nturn), h = nt >> 1;
  int rc;
  if ((rc = check_len(h)) || (rc = check_window(xlow, xhi)) || (rc = check_window(ylow, yhi))) return rc;
  if (total) *total = 0;
  if (np == 0) return ACCP_OK;
  // aline.cpp:1421-1484: each particle from (x0, 0, y0, 0, 0, 0); the tunes
  // of the first and second halves of nt positions.  Particles go in groups
  // whose turn-by-turn record (48 B per particle and turn) stays below
  // RECORD_BYTES; the results do not depend on the grouping.
  const int group = (int)std::max<int64_t>(1, std::min<int64_t>(np, RECORD_BYTES / (48 * (int64_t)nt)));
  const double mdiffu = -4 * std::log(nt * 1.0);
  double tdif = 0;
  for (int g0 = 0; g0 < np; g0 += group) {
    const int n = std::min(group, np - g0);
    std::vector<double> r(6 * (size_t)n, 0.0);
    for (int i = 0; i < n; ++i) {
      r[6 * (size_t)i] = x0[g0 + i];
      r[6 * (size_t)i + 2] = y0[g0 + i];
    }
    const double *rec = nullptr;
    if ((rc = accp::track_record(L, r.data(), n, nt, &rec, (uint64_t)g0))) return rc;
    // lost after nt - 1 turns: x of the last record
    std::vector<double> last(6 * (size_t)n);
    if (hipMemcpy(last.data(), rec + (size_t)(nt - 1) * n * 6, sizeof(double) * 6 * n, hipMemcpyDeviceToHost) !=
        hipSuccess)
      return fail(ACCP_E_HIP, "accp_trackdiffusion: copy");
    std::vector<accp::NaffJob> jobs;
    std::vector<int> live;
    const int64_t stride = 6 * (int64_t)n;
    for (int i = 0; i < n; ++i) {
      if (std::isnan(last[6 * (size_t)i])) continue;
      live.push_back(i);
      for (int half = 0; half < 2; ++half) {
        const int64_t b = ((int64_t)half * h * n + i) * 6;
        jobs.push_back(job(b, stride, h, xlow, xhi));
        jobs.push_back(job(b + 2, stride, h, ylow, yhi));
      }
    }
    std::vector<double> out;
    if ((rc = run_jobs(rec, jobs, out))) return rc;
    size_t k = 0;
    for (int i = 0; i < n; ++i) {
      const int ip = g0 + i;
      if (k < live.size() && live[k] == i) {
        const double nux0 = out[4 * k], nuy0 = out[4 * k + 1], nux1 = out[4 * k + 2], nuy1 = out[4 * k + 3];
        nux[ip] = nux0;
        nuy[ip] = nuy0;
        const double d =
            std::log(std::sqrt((nux1 - nux0) * (nux1 - nux0) + (nuy1 - nuy0) * (nuy1 - nuy0)) * 2.0 / nt);
        dif[ip] = mdiffu > d ? mdiffu : d;
        tdif += dif[ip];
        ++k;
      } else {
        nux[ip] = nuy[ip] = dif[ip] = NAN;
        tdif += -3.0;
      }
    }
  }
  if (total) *total = tdif;
  return ACCP_OK;
}

extern "C" int accp_trackdnudexy(accp_lattice *L, int nturn, double xm, double ym, double dnudexy[4],
                                 double *rows) {
  if (!L || !dnudexy || nturn < 1) return fail(ACCP_E_ARG, "accp_trackdnudexy: bad arguments");
  const int nt = (int)pow2_at_least((unsigned)nturn);
  int rc;
  if ((rc = check_len(nt))) return rc;
  // aline.cpp:716-779: linear tunes of the ring start, then six particles
  // per plane at growing amplitude, nt positions each (before every turn)
  double tw[7];
  if ((rc = accp_calctwiss(L, nullptr, 0, 0.0, ACCP_DELTA_DIFF, tw))) return rc;
  const double tunex = tw[1] / 2 / PI, tuney = tw[2] / 2 / PI;
  const unsigned np = 6;
  const double tiny = 1.0e-5;
  std::vector<double> r(6 * 2 * np, 0.0), de(2 * np);
  for (unsigned k = 0; k < np; ++k) {
    double *a = &r[6 * k], *b = &r[6 * (np + k)];
    a[0] = tiny + k * 1.0 / (np - 1) * xm;
    a[2] = tiny + a[0] * .02;
    de[k] = a[0] * a[0] / tw[3];
    b[2] = tiny + k * 1.0 / (np - 1) * ym;
    b[0] = tiny + b[2] * .02;
    de[np + k] = b[2] * b[2] / tw[5];
  }
  const double xlo = 0 > tunex - 0.1 ? 0 : tunex - 0.1, xhi = 0.5 < tunex + 0.1 ? 0.5 : tunex + 0.1;
  const double ylo = 0 > tuney - 0.1 ? 0 : tuney - 0.1, yhi = 0.5 < tuney + 0.1 ? 0.5 : tuney + 0.1;
  if ((rc = check_window(xlo, xhi)) || (rc = check_window(ylo, yhi))) return rc;
  const double *rec = nullptr;
  if ((rc = accp::track_record(L, r.data(), (int)(2 * np), nt, &rec))) return rc;
  std::vector<accp::NaffJob> jobs;
  for (unsigned p = 0; p < 2 * np; ++p) {
    jobs.push_back(job(6 * (int64_t)p, 6 * 2 * (int64_t)np, nt, xlo, xhi));
    jobs.push_back(job(6 * (int64_t)p + 2, 6 * 2 * (int64_t)np, nt, ylo, yhi));
  }
  std::vector<double> out;
  if ((rc = run_jobs(rec, jobs, out))) return rc;
  for (int plane = 0; plane < 2; ++plane) {
    double xt[6], yt[6];
    for (unsigned k = 0; k < np; ++k) {
      xt[k] = out[2 * (plane * np + k)];
      yt[k] = out[2 * (plane * np + k) + 1];
      if (rows) {
        double *row = rows + 3 * (plane * np + k);
        row[0] = de[plane * np + k];
        row[1] = xt[k];
        row[2] = yt[k];
      }
    }
    dnudexy[2 * plane] = slope((int)np, &de[plane * np], xt);
    dnudexy[2 * plane + 1] = slope((int)np, &de[plane * np], yt);
  }
  return ACCP_OK;
}
