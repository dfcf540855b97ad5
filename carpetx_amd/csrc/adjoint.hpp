// Adjoint (transpose) of the rf=2 prolongation.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cxp.h"
#include "kernels.hpp"
#include "ops.hpp"

namespace cxp {

inline int adjoint_launch(const OpDesc &op, const cxp_fab *crse_in,
                          const cxp_fab *crse_bar, int crse_comp,
                          const cxp_fab *fine_bar, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3], int per_group,
                          hipStream_t s) {
  (void)op; (void)crse_in; (void)crse_bar; (void)crse_comp; (void)fine_bar;
  (void)fine_comp; (void)ncomp; (void)rlo; (void)rhi; (void)per_group; (void)s;
  return CXP_ERR_UNSUPPORTED;
}

} // namespace cxp
