/*
 * cxp_oracle.h -- CPU restatement of CarpetX's rf=2 inter-level transfer.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; it is loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product never calls it.
 *
 * Restates, per fine point, the algorithm of
 *   CarpetX/src/prolongate_3d_rf2_impl.hxx  (1-D stencils :388-878, tensor
 *   product :1178-1216, CoarseBox :1226-1254, per-var :1256-1689, per-group
 *   :1691-2165),
 *   CarpetX/src/driver.cxx:593-642 (operator selection),
 *   CarpetX/src/schedule.cxx:2791-2816 (restriction dispatch to AMReX
 *   average_down_nodal/_edges/_faces/average_down).
 *
 * Parity status per function is listed in the header of cxp_oracle.c.
 */
#ifndef CXP_ORACLE_H
#define CXP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_VC = 0, OR_CC = 1 };
enum { OR_POLY = 0, OR_HERMITE = 1, OR_CONS = 2, OR_ENO = 3, OR_MINMOD = 4 };
enum { OR_FB_NONE = 0, OR_FB_LINEAR = 1 };

typedef struct {
  int cent[3];
  int intp[3];
  int order[3];
  int fb;
} or_op;

/* A FArrayBox view: x fastest, then y, z, then component. */
typedef struct {
  double *data;
  int lo[3];
  int hi[3];
  int64_t comp_stride;
} or_fab;

/* get_interpolator (driver.cxx:593-642). 0 ok; -1 unknown type; -2 order. */
int or_lookup(const char *type, int order, const int indextype[3], or_op *out);

int or_required_ghosts(int cent, int intp, int order);
void or_stencil_radius(int cent, int intp, int order, int shift, int off,
                       int r[2]);

/* CoarseBox (prolongate_3d_rf2_impl.hxx:1226-1254) with AMReX coarsen. */
void or_coarse_box(const or_op *op, const int flo[3], const int fhi[3],
                   int clo[3], int chi[3]);

/* One 1-D operator applied at coarse offset 0 of c (stride st). */
double or_interp1d(int cent, int intp, int order, int shift, int off,
                   const double *c, ptrdiff_t st);

/* Weight table entry (for tests); returns NaN when out of range. */
double or_coeff(int cent, int intp, int order, int row, int i);

/* prolongate_3d_rf2::interp.  0 ok, <0 error (region outside fine FAB,
 * coarse stencil outside crse FAB, ncomp>16 in per-group mode). */
int or_prolongate(const or_op *op, const or_fab *crse, int crse_comp,
                  const or_fab *fine, int fine_comp, int ncomp,
                  const int rlo[3], const int rhi[3], int per_group);

/* Many independent calls, OpenMP across them (as FillPatchInterp iterates
 * boxes).  Returns the first non-zero status. */
typedef struct {
  const or_op *op;
  or_fab crse;
  int crse_comp;
  or_fab fine;
  int fine_comp;
  int ncomp;
  int rlo[3];
  int rhi[3];
} or_item;
int or_prolongate_many(int nitems, const or_item *items, int per_group);

/* Restriction over the coarse region [rlo, rhi]. */
int or_restrict(const int indextype[3], const or_fab *fine, int fine_comp,
                const or_fab *crse, int crse_comp, int ncomp, const int rlo[3],
                const int rhi[3]);

int or_num_threads(void);
/* Threads of the later parallel regions (n > 0). */
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
