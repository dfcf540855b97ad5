/*
 * cxp_wave.h -- the native RK4 stage of the two-level wave harness
 * (carpetx_amd/libcxp_wave.so).
 *
 * No counterpart in the reference's transfer path: the harness evolves the
 * scalar wave equation u_t = rho, rho_t = Laplace u on the two-level grid
 * whose ghost fills and restriction go through cxp.h, as an RK4 ODESolvers
 * step does around CarpetX's SYNCs (ODESolvers/src/solve.cxx:888-916).
 * This entry point is its right-hand side and stage update, fused.
 */
#ifndef CXP_WAVE_H
#define CXP_WAVE_H

#include "cxp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One stage (1..4) on a box of nx x ny x nz points and 2 components (u,
 * rho; x fastest, component stride nx ny nz), at the interior points
 * [ghosts, n - ghosts) per direction (periodic: ghosts = 0, neighbours
 * wrap).  k = (rho, c L u) with L the 4th-order centred Laplacian
 * (-u[-2] + 16 u[-1] - 30 u + 16 u[+1] - u[+2] per direction, c =
 * 1 / (12 h^2)); the Laplacian reads two points each way, so a
 * non-periodic box needs ghosts >= 2 (CXP_ERR_ARG otherwise); then
 *   stage 1:    acc = k;          dst = base + a k
 *   stage 2, 3: acc = acc + 2 k;  dst = base + a k
 *   stage 4:                      dst = base + a (acc + k)
 * at the interior points.  y must not be dst; base may be dst.  Device
 * pointers, asynchronous on the stream. */
int cxp_wave_rk4_stage(int stage, const double *y, double *dst,
                       const double *base, double *acc, int nx, int ny,
                       int nz, int ghosts, int periodic, double c, double a,
                       void *stream);

/* Adjoint of cxp_wave_rk4_stage with respect to y, in gather form:
 * y_bar += (dk / dy)^T k_bar with k_bar = a dst_bar + w acc_bar at the
 * interior points (w = 1 for stage 1, 2 for stages 2 and 3; stage 4: k_bar
 * = a dst_bar, acc_bar unused).  y_bar covers the whole box (ghosts
 * included); the gradients of base and of the incoming acc are pointwise
 * and left to the caller.  The same argument rules as the forward. */
int cxp_wave_rk4_stage_adjoint(int stage, const double *dst_bar,
                               const double *acc_bar, double *y_bar, int nx,
                               int ny, int nz, int ghosts, int periodic,
                               double c, double a, void *stream);

#ifdef __cplusplus
}
#endif

#endif
