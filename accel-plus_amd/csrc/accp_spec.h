// accp_spec.h -- the lattice-specialized tracker: the persistent tracker of
// accp_track.h compiled at run time (hiprtc) for one lattice structure, with
// the turn's element sequence (or a range of it) written out as straight-line
// code instead of a walk over the element table.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "accp_ctx.h"
#include "accp_table.h"

namespace accp {

// Groups [s, s + p*r) of a table are r copies of the p groups at s: equal
// op, a and b, and constants `stride` doubles after the previous copy's.
struct Period {
  int s = 0, p = 0, r = 0, stride = 0;
};

// The repeated block covering the most groups (r >= 2, p <= max_p; the
// shortest period among equals), or r = 0 when there is none.
Period find_period(const std::vector<Hdr> &h, int max_p);

// Which entries [st, ed) of a table a specialized kernel runs, and in which
// tracker variant: full = the element table with every element map (radiation,
// noise, linear Bend, ...), otherwise the grouped table of the lean variant.
struct SpecRange {
  int st = 0, ed = 0;
  bool full = false;
};

// Why entries [r.st, r.ed) of table h cannot be specialized, or "" when they
// can (ops of the variant, code within a bounded number of straight-line
// entries).
std::string spec_unsupported(const std::vector<Hdr> &h, const SpecRange &r);

// HIP source of the specialized kernel `accp_spec_track` for entries
// [r.st, r.ed) of table h, launch shape (block, ppl) and staging (lds).
std::string spec_source(const std::vector<Hdr> &h, const SpecRange &r, int block, int ppl, bool lds);

// Compiles source (from spec_source) for offload target arch ("gfx950",
// or a device's gcnArchName) into a code object: ACCP_OK, or an error code
// with the compiler's log.  No device needed.
int spec_compile(const std::string &src, const std::string &arch, std::vector<char> &code, std::string &log);

struct SpecKernel;

// The compiled kernel for (h, r, block, ppl, lds) on the current device,
// compiled on first use and kept for the life of the process.  Call it
// outside any stream capture (module loading is not a stream operation).
int spec_kernel(const std::vector<Hdr> &h, const SpecRange &r, int block, int ppl, bool lds, const SpecKernel **out);

hipError_t launch_spec(const SpecKernel *k, const KernelCtx &ctx, int grid, size_t lds_bytes, hipStream_t s);

}  // namespace accp
