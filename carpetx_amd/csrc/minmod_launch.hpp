// Host launches of the MINMOD adjoint kernels (minmod_adjoint.hip): the
// branch record, the gather from recorded branches and the fused CCC
// adjoint.  Each returns a CXP_* status; arguments are those the C ABI has
// already checked.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cxp.h"
#include "ops.hpp"

namespace cxp {

inline bool op_has_minmod(const OpDesc &op) {
  for (int d = 0; d < 3; ++d)
    if (op.sid[d] == SID_MINMOD)
      return true;
  return false;
}

inline bool op_all_minmod(const OpDesc &op) {
  return op.sid[0] == SID_MINMOD && op.sid[1] == SID_MINMOD &&
         op.sid[2] == SID_MINMOD;
}

// The tiled CCC kernels address a plane of fine_bar with 32-bit offsets.
inline bool mm_tiled_fits(const cxp_fab *fine_bar) {
  return (long long)(fine_bar->hi[0] - fine_bar->lo[0] + 1) *
             (fine_bar->hi[1] - fine_bar->lo[1] + 1) <
         (1LL << 31);
}

// Bytes of branch codes of a MINMOD operator over the region: one per point
// each MINMOD pass reads along its direction, per component.
long long mm_code_bytes(const OpDesc &op, const int rlo[3], const int rhi[3],
                        int ncomp);

// Records the branch codes of every component from the forward input.
int mm_record_launch(const OpDesc &op, const cxp_fab *crse_in, int crse_comp,
                     int ncomp, const int rlo[3], const int rhi[3],
                     unsigned char *codes, hipStream_t s);

// crse_bar += P^T fine_bar from recorded codes (tiled for CCC operators,
// one thread per coarse point of the output window else).
int mm_gather_launch(const OpDesc &op, const cxp_fab *crse_bar, int crse_comp,
                     const cxp_fab *fine_bar, int fine_comp, int ncomp,
                     const int rlo[3], const int rhi[3],
                     const unsigned char *codes, hipStream_t s);

// crse_bar += J(crse_in)^T fine_bar for a CCC MINMOD operator in one
// launch that recomputes the branches (no code array); needs
// mm_tiled_fits(fine_bar).
int mm_fused_launch(const OpDesc &op, const cxp_fab *crse_in,
                    const cxp_fab *crse_bar, int crse_comp,
                    const cxp_fab *fine_bar, int fine_comp, int ncomp,
                    const int rlo[3], const int rhi[3], hipStream_t s);

} // namespace cxp
