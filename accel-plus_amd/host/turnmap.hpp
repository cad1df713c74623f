// turnmap.hpp -- the turn map as a differentiable function: forward tracks a
// beam through nturn turns keeping checkpoints, backward takes dL/d(final
// coordinates) and returns dL/d(knobs) and dL/d(initial coordinates).
//
// A C++ host-side wrapper of the accp_adjoint_* entry points of accp.h.  It
// replaces the finite-difference responses of ALine::calctuneresp and
// calcchromresp (src/lib/accel/aline.cpp:1187-1242, :1310-1355): one forward
// and one backward give the derivative with respect to every knob at once.
// Errors are reported as std::runtime_error carrying accp_last_error().
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/accp.h"

namespace accp {
namespace host {

// One knob: every element of `members` moves together (a family shift).
struct Knob {
  int field;  // ACCP_FIELD_K, ACCP_FIELD_POLYB or ACCP_FIELD_POLYA
  int index;  // multipole index for PolyB/PolyA
  std::vector<int32_t> members;
};

struct Gradient {
  std::vector<double> dparam;  // dL/d(knob), one per knob
  std::vector<double> dr0;     // dL/d(initial coordinates), AoS like r0
};

class TurnMapFunction {
 public:
  TurnMapFunction(accp_lattice *lat, const std::vector<Knob> &knobs) : knobs_(knobs) {
    std::vector<accp_param> p(knobs_.size());
    for (size_t k = 0; k < knobs_.size(); ++k)
      p[k] = accp_param{knobs_[k].field, knobs_[k].index, (int32_t)knobs_[k].members.size(),
                        knobs_[k].members.data()};
    check(accp_adjoint_create(lat, (int)p.size(), p.data(), &adj_), "accp_adjoint_create");
  }
  ~TurnMapFunction() { accp_adjoint_free(adj_); }
  TurnMapFunction(const TurnMapFunction &) = delete;
  TurnMapFunction &operator=(const TurnMapFunction &) = delete;

  // Final coordinates of r0 (AoS, 6 per particle) after nturn turns.
  std::vector<double> forward(const std::vector<double> &r0, int nturn, int ckpt_every = 16) {
    if (r0.size() % 6) throw std::invalid_argument("r0 must hold 6 coordinates per particle");
    std::vector<double> out(r0.size());
    check(accp_adjoint_forward(adj_, r0.data(), (int)(r0.size() / 6), nturn, ckpt_every, out.data()),
          "accp_adjoint_forward");
    n_ = r0.size();
    return out;
  }

  // dL/dknobs and dL/dr0 for the given dL/d(final coordinates); may be
  // called repeatedly after one forward (one call per scalar loss).
  Gradient backward(const std::vector<double> &dl_dr_out) {
    if (dl_dr_out.size() != n_) throw std::invalid_argument("dl_dr_out does not match the forward beam");
    Gradient g{std::vector<double>(knobs_.size()), std::vector<double>(n_)};
    check(accp_adjoint_backward(adj_, dl_dr_out.data(), g.dparam.data(), g.dr0.data()),
          "accp_adjoint_backward");
    return g;
  }

 private:
  static void check(int rc, const char *what) {
    if (rc != ACCP_OK) throw std::runtime_error(std::string(what) + ": " + accp_last_error());
  }
  std::vector<Knob> knobs_;
  accp_adjoint *adj_ = nullptr;
  size_t n_ = 0;
};

}  // namespace host
}  // namespace accp
