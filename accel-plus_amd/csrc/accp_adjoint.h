// accp_adjoint.h -- reverse-mode turn map (declarations shared by the C ABI
// layer and the adjoint kernels).
#pragma once
