// Host launches of the CCC ENO kernels (eno_ccc.hip).  Arguments are those
// the C ABI has already checked.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cxp.h"
#include "ops.hpp"

namespace cxp {

// ENO in all three directions, per-var, no linear fallback.
bool eno_ccc_fits(const OpDesc &op, int per_group);

// The forward over [rlo, rhi]: values into fine (or, with codes_only, no
// values and fine may be null), and each fine point's branch code into
// codes when given (per component, region-linear, x fastest).  With
// codes_only and parent_codes: one code per coarse parent instead
// (per component, parent-linear over [rlo >> 1, rhi >> 1]).
int eno_ccc_launch(const cxp_fab *crse, int crse_comp, const cxp_fab *fine,
                   int fine_comp, int ncomp, const int rlo[3],
                   const int rhi[3], unsigned char *codes, bool codes_only,
                   hipStream_t s, bool parent_codes = false);

// crse_bar += J^T fine_bar from the branch codes of eno_ccc_launch (or of
// the generic record: the same codes), in the layout parent_codes names.
int eno_ccc_adj_launch(const cxp_fab *crse_bar, int crse_comp,
                       const cxp_fab *fine_bar, int fine_comp, int ncomp,
                       const int rlo[3], const int rhi[3],
                       const unsigned char *codes, hipStream_t s,
                       bool parent_codes = false);

} // namespace cxp
