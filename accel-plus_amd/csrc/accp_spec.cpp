// accp_spec.cpp -- the lattice-specialized tracker, compiled at run time.
//
// The table-walking tracker dispatches every group of a turn on its header:
// a chain of scalar compares and branches, the group's constants read from
// LDS behind it, and the state copied between the branches' registers.  When
// a turn is one wave's instruction stream (the sparse tail of a DA scan: a
// few surviving particles per seed, one wave each) that dispatch is a third
// of the turn.  Here the host writes the turn out as a sequence of calls
// with literal headers, hiprtc compiles it with the same device code
// (accp_track.h), and the compiler folds every dispatch away: only the
// element maps and the constants' loads (at fixed LDS offsets, free to move
// ahead of their use) remain.  A ring of identical cells becomes a loop over
// one cell, so the code stays within the instruction cache.
//
// Host code (HIP runtime, hiprtc).
#include "accp_spec.h"

#include <hip/hiprtc.h>

#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <utility>

#include "../../include/accp.h"
#include "accp_error.h"

// the device headers the kernel compiles against, embedded at build time
extern const char accp_src_accp_table[], accp_src_accp_ctx[], accp_src_accp_physics[], accp_src_accp_track[];

namespace accp {

namespace {

constexpr int MAX_PERIOD = 64;     // entries in one loop body
constexpr int MAX_STRAIGHT = 256;  // entries written out once (outside the loop, plus its body)

bool same_shape(const Hdr &a, const Hdr &b) { return a.op == b.op && a.a == b.a && a.b == b.b; }

// Ops the lean tracker runs (the specialized kernel is the lean variant).
bool lean_op(int32_t op) {
  switch (op & FUSE_OP_MASK) {
    case OP_NOP:
    case OP_DRIFT:
    case OP_QUAD:
    case OP_QUAD_K0:
    case OP_BENDMP:
    case OP_STRMP:
    case OP_STRMP_THIN:
    case OP_CAVITY:
    case OP_APERTURE:
    case OP_CORRECTOR:
    case OP_WAKE:  // the host splits the turn at an active wake; inside a pass it does nothing
      return true;
    default:
      return false;
  }
}

std::mutex g_mu;
// Compiled kernels by (device, source), and by (device, the entries' raw
// headers, range and launch shape): a call that finds its key in the second
// map builds no source (a turn split at a wake asks once per segment and turn).
std::map<std::pair<int, std::string>, SpecKernel *> g_cache;
std::map<std::pair<int, std::string>, SpecKernel *> g_by_table;

std::string table_key(const std::vector<Hdr> &h, const SpecRange &r, int block, int ppl, bool lds) {
  const int32_t shape[6] = {r.st, r.ed, r.full ? 1 : 0, block, ppl, lds ? 1 : 0};
  std::string k(reinterpret_cast<const char *>(shape), sizeof shape);
  k.append(reinterpret_cast<const char *>(h.data() + r.st), sizeof(Hdr) * (size_t)(r.ed - r.st));
  return k;
}

int err(int code, const std::string &msg) { return set_error(code, msg.c_str()); }

}  // namespace

struct SpecKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  int block = 0;
};

Period find_period(const std::vector<Hdr> &h, int max_p) {
  const int g = (int)h.size();
  Period best;
  for (int p = 1; p <= max_p && 2 * p <= g; ++p) {
    // runs of i where group i + p repeats group i with one constant stride
    int i = 0;
    while (i + p < g) {
      if (!same_shape(h[i], h[i + p])) {
        ++i;
        continue;
      }
      const int d = h[i + p].off - h[i].off;
      int j = i + 1;
      while (j + p < g && same_shape(h[j], h[j + p]) && h[j + p].off - h[j].off == d) ++j;
      const int run = j - i;  // groups i .. j-1 each repeat p later
      const int r = run / p + 1;
      if (r >= 2 && (int64_t)p * r > (int64_t)best.p * best.r) best = Period{i, p, r, d};
      i = j;
    }
  }
  return best;
}

std::string spec_unsupported(const std::vector<Hdr> &h, const SpecRange &r) {
  if (r.st < 0 || r.ed > (int)h.size() || r.st >= r.ed) return "empty range";
  const std::vector<Hdr> sub(h.begin() + r.st, h.begin() + r.ed);
  if (!r.full)
    for (const Hdr &x : sub)
      if (!lean_op(x.op)) return "an element needs the full tracker variant";
  const Period pr = find_period(sub, MAX_PERIOD);
  const int straight = (int)sub.size() - (pr.r >= 2 ? pr.p * (pr.r - 1) : 0);
  if (straight > MAX_STRAIGHT) return "more than 256 elements outside a repeated cell";
  return "";
}

std::string spec_source(const std::vector<Hdr> &hall, const SpecRange &r, int block, int ppl, bool lds) {
  const std::vector<Hdr> h(hall.begin() + r.st, hall.begin() + r.ed);
  const int g = (int)h.size();
  const Period pr = find_period(h, MAX_PERIOD);
  const char *full = r.full ? "true" : "false";
  std::ostringstream o;
  o << "// the persistent tracker specialized for one lattice structure: entries " << r.st << ".." << r.ed - 1
    << " of the " << (r.full ? "element" : "grouped") << " table";
  if (pr.r >= 2)
    o << ", entries " << r.st + pr.s << ".." << r.st + pr.s + pr.p * pr.r - 1 << " as " << pr.r << " cells of "
      << pr.p;
  o << "\n#include \"accp_track.h\"\n\nnamespace accp {\n\n"
       "struct SpecWalker {\n"
       "  template <int PPL, bool FULL, bool LDS, class PT>\n"
       "  __device__ static __forceinline__ void turn(Part (&p)[PPL], const Hdr *, const PT *par, const KernelCtx &ctx,\n"
       "                                              const uint64_t (&pid)[PPL], uint64_t turn,\n"
       "                                              const int64_t (&gidx)[PPL], bool (&parked)[PPL]) {\n";
  auto emit = [&](const Hdr &x, const char *base, int off, const std::string &e, const char *indent) {
    o << indent << "track_step<PPL, FULL>(p, Hdr{" << x.op << ", " << off << ", " << x.a << ", " << x.b << "}, "
      << base << " + " << off << ", ctx, pid, turn, " << e << ", gidx, parked);\n";
  };
  int i = 0;
  if (pr.r >= 2) {
    for (; i < pr.s; ++i) emit(h[i], "par", h[i].off, std::to_string(r.st + i), "    ");
    o << "#pragma unroll 1\n"
      << "    for (int c = 0; c < " << pr.r << "; ++c) {\n"
      << "      const PT *q = par + " << h[pr.s].off << " + c * " << pr.stride << ";\n"
      << "      const int e0 = " << r.st + pr.s << " + c * " << pr.p << ";\n";
    for (int j = 0; j < pr.p; ++j)
      emit(h[pr.s + j], "q", h[pr.s + j].off - h[pr.s].off, "e0 + " + std::to_string(j), "      ");
    o << "    }\n";
    i = pr.s + pr.p * pr.r;
  }
  for (; i < g; ++i) emit(h[i], "par", h[i].off, std::to_string(r.st + i), "    ");
  o << "  }\n};\n\n}  // namespace accp\n\n"
    << "extern \"C\" __global__ __launch_bounds__(" << block << ") void accp_spec_track(accp::KernelCtx ctx) {\n"
    << "  accp::track_tile<" << block << ", " << ppl << ", " << (lds ? "true" : "false") << ", " << full
    << ", accp::SpecWalker>(ctx);\n}\n";
  return o.str();
}

int spec_compile(const std::string &src, const std::string &arch, std::vector<char> &code, std::string &log) {
  code.clear();
  log.clear();
  const std::string target = "--offload-arch=" + arch;
  const char *hdrs[4] = {accp_src_accp_table, accp_src_accp_ctx, accp_src_accp_physics, accp_src_accp_track};
  const char *names[4] = {"accp_table.h", "accp_ctx.h", "accp_physics.h", "accp_track.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "accp_spec.hip", 4, hdrs, names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return ACCP_E_HIP;
  }
  // the flags of the prebuilt kernels: -O3, packed FP32 operations off
  const char *opts[] = {target.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t n = 0;
  if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
    log.resize(n);
    (void)hiprtcGetProgramLog(prog, &log[0]);
    log.resize(strnlen(log.c_str(), n));
  }
  if (rc != HIPRTC_SUCCESS) {
    (void)hiprtcDestroyProgram(&prog);
    log = std::string("compilation failed: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 2000);
    return ACCP_E_HIP;
  }
  size_t nc = 0;
  if (hiprtcGetCodeSize(prog, &nc) == HIPRTC_SUCCESS && nc > 0) {
    code.resize(nc);
    if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
  }
  (void)hiprtcDestroyProgram(&prog);
  if (code.empty()) {
    log = "compilation produced no code object";
    return ACCP_E_HIP;
  }
  return ACCP_OK;
}

int spec_kernel(const std::vector<Hdr> &h, const SpecRange &r, int block, int ppl, bool lds,
                const SpecKernel **out) {
  *out = nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return err(ACCP_E_HIP, "specialized tracker: no current HIP device");
  if (r.st < 0 || r.ed > (int)h.size() || r.st >= r.ed) return err(ACCP_E_ARG, "specialized tracker: empty range");
  const std::string tkey = table_key(h, r, block, ppl, lds);
  std::lock_guard<std::mutex> lock(g_mu);
  auto hit = g_by_table.find({dev, tkey});
  if (hit != g_by_table.end()) {
    *out = hit->second;
    return ACCP_OK;
  }
  const std::string why = spec_unsupported(h, r);
  if (!why.empty()) return err(ACCP_E_ARG, "the lattice cannot be specialized: " + why);
  const std::string src = spec_source(h, r, block, ppl, lds);
  auto it = g_cache.find({dev, src});
  if (it != g_cache.end()) {
    g_by_table[{dev, tkey}] = it->second;
    *out = it->second;
    return ACCP_OK;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return err(ACCP_E_HIP, "specialized tracker: hipGetDeviceProperties failed");
  std::vector<char> code;
  std::string log;
  int rc = spec_compile(src, prop.gcnArchName, code, log);
  if (rc) return err(rc, "specialized tracker: " + log);
  SpecKernel *k = new SpecKernel;
  k->block = block;
  hipError_t e = hipModuleLoadData(&k->mod, code.data());
  if (e == hipSuccess) e = hipModuleGetFunction(&k->fn, k->mod, "accp_spec_track");
  if (e != hipSuccess) {
    if (k->mod) (void)hipModuleUnload(k->mod);
    delete k;
    return err(ACCP_E_HIP, std::string("specialized tracker: module load failed: ") + hipGetErrorString(e));
  }
  g_cache[{dev, src}] = k;
  g_by_table[{dev, tkey}] = k;
  *out = k;
  return ACCP_OK;
}

hipError_t launch_spec(const SpecKernel *k, const KernelCtx &ctx, int grid, size_t lds_bytes, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  KernelCtx c = ctx;
  void *args[] = {&c};
  return hipModuleLaunchKernel(k->fn, (unsigned)grid, 1, 1, (unsigned)k->block, 1, 1, (unsigned)lds_bytes, s, args,
                               nullptr);
}

}  // namespace accp
