/*
 * cxp_ref.h -- the CPU oracle behind the C interface's own signatures.
 *
 * TEST INFRASTRUCTURE ONLY (see cxp_oracle.h): the checker and CPU baseline,
 * never the product.  These take the same arguments as cxp_op_lookup,
 * cxp_coarse_box, cxp_prolongate and cxp_restrict in include/cxp.h, with
 * host pointers in the cxp_fab structs, and return the same status codes
 * for the same argument errors, so a caller can run one call through both
 * libraries and compare.  The computation is the oracle's restatement of
 * CarpetX/src/prolongate_3d_rf2_impl.hxx and of the restriction dispatch
 * of CarpetX/src/schedule.cxx:2791-2816, synchronous on the calling thread
 * (the stream argument is ignored).  Differences from the device library:
 * operator handles of the two libraries are not interchangeable, and there
 * is no 2 GiB limit per component.
 */
#ifndef CXP_REF_H
#define CXP_REF_H

#include "../include/cxp.h"

#ifdef __cplusplus
extern "C" {
#endif

int cxp_ref_op_lookup(const char *type, int order, const int indextype[3],
                      cxp_op *out);
int cxp_ref_coarse_box(cxp_op op, const int fine_lo[3], const int fine_hi[3],
                       int crse_lo[3], int crse_hi[3]);
int cxp_ref_prolongate(cxp_op op, const cxp_fab *crse, int crse_comp,
                       const cxp_fab *fine, int fine_comp, int ncomp,
                       const int region_lo[3], const int region_hi[3],
                       int per_group, void *stream);
int cxp_ref_restrict(const int indextype[3], const cxp_fab *fine,
                     int fine_comp, const cxp_fab *crse, int crse_comp,
                     int ncomp, const int crse_region_lo[3],
                     const int crse_region_hi[3], void *stream);

#ifdef __cplusplus
}
#endif

#endif
