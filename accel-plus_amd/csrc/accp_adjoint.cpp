// accp_adjoint.cpp -- C ABI of the reverse-mode turn map.
#include <hip/hip_runtime.h>

#include "../../include/accp.h"

extern "C" int accp_adjoint_create(accp_lattice *, int, const accp_param *, accp_adjoint **out) {
  if (out) *out = nullptr;
  return ACCP_E_STATE;
}
extern "C" int accp_adjoint_forward(accp_adjoint *, const double *, int, int, int, double *) {
  return ACCP_E_STATE;
}
extern "C" int accp_adjoint_backward(accp_adjoint *, const double *, double *, double *) {
  return ACCP_E_STATE;
}
extern "C" int accp_adjoint_free(accp_adjoint *) { return ACCP_OK; }
