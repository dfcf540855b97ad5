/*
 * cxp.h -- C ABI of the MI355X-native rf=2 AMR inter-level transfer library
 * (libcxp.so).  Plain C: device pointers, sizes and status codes; streams
 * are hipStream_t passed as void* (NULL = the default stream).
 *
 * Each entry point names the reference interface it replaces.  All device
 * work is enqueued on the given stream and the call returns without
 * synchronising, like the reference's ParallelFor launches
 * (CarpetX/src/prolongate_3d_rf2_impl.hxx:1685-1688).
 *
 * Data layout: a FAB is AMReX's FArrayBox layout -- x fastest, then y, z,
 * then component, indices in the level's global index space, boxes inclusive
 * [lo, hi] (CarpetX/src/prolongate_3d_rf2_impl.hxx:1339-1348,1776-1779).
 */
#ifndef CXP_H
#define CXP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference aborts (assert / CCTK_VERROR) where these are
 * returned. */
#define CXP_OK 0
#define CXP_ERR_TYPE -1        /* unknown prolongation_type (driver.cxx:620) */
#define CXP_ERR_ORDER -2       /* unsupported order for type (driver.cxx:623) */
#define CXP_ERR_ARG -3         /* null pointer, bad indextype, ncomp < 1 ... */
#define CXP_ERR_REGION -4      /* fine_region not inside fine FAB (:1300-1302) */
#define CXP_ERR_COARSE -5      /* crse FAB does not cover CoarseBox(region) */
#define CXP_ERR_NCOMP -6       /* per-group mode with ncomp > 16 (:1707-1708) */
#define CXP_ERR_RATIO -7       /* refinement ratio other than 2 (:1293-1294) */
#define CXP_ERR_HIP -8         /* a HIP runtime call failed */
#define CXP_ERR_COMP -9        /* component range outside the FAB */
#define CXP_ERR_STATE -10      /* worklist / graph used in the wrong state */
#define CXP_ERR_UNSUPPORTED -11 /* operator has no adjoint of this kind */
#define CXP_ERR_SCRATCH -12    /* exceeds the no-allocation scratch array */

/* Centring and interpolation codes (prolongate_3d_rf2.hxx:15-31). */
#define CXP_VC 0
#define CXP_CC 1
#define CXP_POLY 0
#define CXP_HERMITE 1
#define CXP_CONS 2
#define CXP_ENO 3
#define CXP_MINMOD 4

typedef struct cxp_fab {
  double *data;        /* device pointer to component 0 at (lo) */
  int lo[3];
  int hi[3];
  int64_t comp_stride; /* elements between components (AMReX: box volume) */
  int ncomp;           /* number of components stored */
} cxp_fab;

typedef const struct cxp_op_s *cxp_op; /* immutable, thread-safe, never freed */

int cxp_version(void);
const char *cxp_status_string(int status);

/* get_interpolator(type, order, indextype) -- CarpetX/src/driver.cxx:593-642.
 * indextype[d]: 0 vertex (VC), 1 cell (CC). */
int cxp_op_lookup(const char *type, int order, const int indextype[3],
                  cxp_op *out);

/* cxp_op_lookup with the linear fallback chosen by the caller.  No
 * reference counterpart: there the fallback is part of each family's
 * operator type (prolongate_3d_rf2.hxx:33-62; FB_LINEAR for poly-cons3lfb
 * and poly-eno3lfb).  fallback CXP_FB_KEEP keeps the family's, CXP_FB_NONE
 * or CXP_FB_LINEAR replaces it; the fallback is active where the reference
 * makes it active (CONS / ENO directions above order 1 per variable, CONS
 * directions per group, impl.hxx:1535-1538, 1990-1992). */
#define CXP_FB_KEEP -1
#define CXP_FB_NONE 0
#define CXP_FB_LINEAR 1
int cxp_op_lookup_fallback(const char *type, int order,
                           const int indextype[3], int fallback,
                           cxp_op *out);

/* get_interpolator(groupdata) -- CarpetX/src/driver.cxx:558-591, 628-642:
 * the operator of a grid-function group whose tags (the TAGS string of its
 * interface.ccl declaration, key=value pairs; may be NULL) may set
 * prolongation_type="..." and prolongation_order=N; unset keys take the
 * driver parameters default_type and default_order.  Keys compare
 * case-insensitively.  Errors: as cxp_op_lookup; CXP_ERR_ARG for tags that
 * do not parse or a prolongation_order that is not an integer (the
 * reference asserts). */
int cxp_group_op_lookup(const char *tags, const char *default_type,
                        int default_order, const int indextype[3],
                        cxp_op *out);

/* Per-direction description: centring, interpolation, order, required
 * ghosts; fb = 1 for the FB_LINEAR families
 * (prolongate_3d_rf2.hxx:64-75 indextype()/interpolation()/order()). */
int cxp_op_describe(cxp_op op, int cent[3], int intp[3], int order[3],
                    int ghosts[3], int *fb);

/* prolongate_3d_rf2::CoarseBox(fine, ratio=2) -- impl.hxx:1226-1254. */
int cxp_coarse_box(cxp_op op, const int fine_lo[3], const int fine_hi[3],
                   int crse_lo[3], int crse_hi[3]);

/* prolongate_3d_rf2::interp(crse, crse_comp, fine, fine_comp, ncomp,
 * fine_region, ratio, ...) -- impl.hxx:2167-2188 (per_group selects
 * interp_per_group, the `prolongate_per_group` parameter).  Writes exactly
 * fine_region of components [fine_comp, fine_comp+ncomp); reads only
 * CoarseBox(fine_region) of crse. */
int cxp_prolongate(cxp_op op, const cxp_fab *crse, int crse_comp,
                   const cxp_fab *fine, int fine_comp, int ncomp,
                   const int region_lo[3], const int region_hi[3],
                   int per_group, void *stream);

/* cxp_prolongate over several regions of one (crse, fine) FAB pair, such as
 * the six slabs of a box's ghost shell: region i is [region_lo[3 i ..
 * 3 i + 2], region_hi[3 i .. 3 i + 2]].  No reference counterpart (there,
 * interp is called once per region); the results are cxp_prolongate's,
 * region by region, bit for bit.  For a separable operator and at most 8
 * regions it is one launch.  Errors: those of cxp_prolongate for any
 * region, CXP_ERR_ARG for a negative count or null arrays, CXP_ERR_REGION
 * when two regions share a point. */
int cxp_prolongate_regions(cxp_op op, const cxp_fab *crse, int crse_comp,
                           const cxp_fab *fine, int fine_comp, int ncomp,
                           int nregions, const int *region_lo,
                           const int *region_hi, int per_group, void *stream);

/* cxp_prolongate_adjoint over several regions of one FAB pair (the
 * transpose of cxp_prolongate_regions): crse_bar += P^T fine_bar for each
 * region in turn, for operators without branches (POLY, HERMITE, CONS;
 * CXP_ERR_UNSUPPORTED for MINMOD, ENO and linear-fallback operators, whose
 * adjoints need the forward input or its branch record).  No reference
 * counterpart.  Regions run in order; consecutive regions whose coarse
 * output windows are apart share a launch when the operator has one
 * stencil in every direction, so the sums are the region-by-region calls'
 * bit for bit.  Errors: those of cxp_prolongate_adjoint for any region,
 * CXP_ERR_ARG for a negative count or null arrays. */
int cxp_prolongate_adjoint_regions(cxp_op op, const cxp_fab *crse_bar,
                                   int crse_comp, const cxp_fab *fine_bar,
                                   int fine_comp, int ncomp, int nregions,
                                   const int *region_lo, const int *region_hi,
                                   int per_group, void *stream);

/* Restrict() -> AMReX average_down_nodal/_edges/_faces/average_down by the
 * number of cell-centred directions (CarpetX/src/schedule.cxx:2791-2816):
 * injection in VC directions, mean of the two children in CC directions,
 * over the coarse region [crse_lo, crse_hi]. */
int cxp_restrict(const int indextype[3], const cxp_fab *fine, int fine_comp,
                 const cxp_fab *crse, int crse_comp, int ncomp,
                 const int crse_region_lo[3], const int crse_region_hi[3],
                 void *stream);

/* Adjoint of cxp_prolongate (no reference counterpart; the reference has no
 * adjoints): crse_bar += P^T fine_bar over CoarseBox(region), in gather form
 * over coarse points, no atomics.  Linear operators (POLY, HERMITE, CONS)
 * ignore crse_in.  ENO and linear-fallback operators are linearised at the
 * forward input crse_in (required; CXP_ERR_ARG without it): each fine
 * point's branch (ENO shift, fallback decision) is recomputed from crse_in
 * with the forward's decision code, and the transpose of that branch's
 * stencil is applied; per_group selects the per-group decisions.  MINMOD
 * operators (impl.hxx:846-878) are linearised at crse_in too (required):
 * each MINMOD application of the x, y and z passes takes one of its three
 * branches (no slope, the minus or the plus difference), recorded per point
 * the pass reads, and the chain of the branches' weights is gathered per
 * coarse point.  The branch codes of both kinds live in a buffer of the
 * calling stream (per device), grown on demand with stream-ordered
 * allocation and reused by later calls on that stream; so these calls
 * cannot be captured in a graph (CXP_ERR_STATE on a capturing stream):
 * capture cxp_prolongate_adjoint_recorded with a caller-owned branch
 * record, or cxp_prolongate_adjoint_noalloc, instead. */
int cxp_prolongate_adjoint(cxp_op op, const cxp_fab *crse_in,
                           const cxp_fab *crse_bar, int crse_comp,
                           const cxp_fab *fine_bar, int fine_comp, int ncomp,
                           const int region_lo[3], const int region_hi[3],
                           int per_group, void *stream);

/* cxp_prolongate_adjoint that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results (bit for bit).  The ENO,
 * linear-fallback and MINMOD operators keep their branch codes
 * (cxp_branch_bytes of the call) in a static device array of
 * CXP_NOALLOC_SCRATCH_BYTES bytes,
 * one per device, held until the process ends; a larger call returns
 * CXP_ERR_SCRATCH before any device work.  The other operators use no
 * scratch.  Every call on one device shares that device's array, so calls
 * that overlap on one device (on two streams, or a graph replay beside
 * another call) race; the array keeps the last call's codes, which every
 * call rewrites in full before reading. */
#define CXP_NOALLOC_SCRATCH_BYTES (16 << 20)
int cxp_prolongate_adjoint_noalloc(cxp_op op, const cxp_fab *crse_in,
                                   const cxp_fab *crse_bar, int crse_comp,
                                   const cxp_fab *fine_bar, int fine_comp,
                                   int ncomp, const int region_lo[3],
                                   const int region_hi[3], int per_group,
                                   void *stream);

/* Bytes of the branch record of a forward over [region_lo, region_hi]
 * (no reference counterpart): one byte per fine point (per component in
 * per-var mode) for ENO and linear-fallback operators; for MINMOD
 * operators one byte per point each MINMOD pass reads (about 7/8 per fine
 * point for CCC), per component; 0 for the linear operators. */
int cxp_branch_bytes(cxp_op op, const int region_lo[3],
                     const int region_hi[3], int ncomp, int per_group,
                     int64_t *bytes);
/* cxp_prolongate that also writes each fine point's branch (ENO shift,
 * fallback decision) into branches, device memory of cxp_branch_bytes bytes
 * (no reference counterpart).  Operators without branches ignore branches
 * (may be null); for the others it is required (CXP_ERR_ARG). */
int cxp_prolongate_record(cxp_op op, const cxp_fab *crse, int crse_comp,
                          const cxp_fab *fine, int fine_comp, int ncomp,
                          const int region_lo[3], const int region_hi[3],
                          int per_group, unsigned char *branches,
                          void *stream);
/* cxp_prolongate_adjoint from a branch record of cxp_prolongate_record over
 * the same region, components and mode (no reference counterpart): the
 * gather alone, without recomputing the branches from crse_in.  Linear
 * operators ignore branches. */
int cxp_prolongate_adjoint_recorded(cxp_op op, const unsigned char *branches,
                                    const cxp_fab *crse_bar, int crse_comp,
                                    const cxp_fab *fine_bar, int fine_comp,
                                    int ncomp, const int region_lo[3],
                                    const int region_hi[3], int per_group,
                                    void *stream);

/* Adjoint of cxp_restrict: fine_bar += R^T crse_bar. */
int cxp_restrict_adjoint(const int indextype[3], const cxp_fab *crse_bar,
                         int crse_comp, const cxp_fab *fine_bar, int fine_comp,
                         int ncomp, const int crse_region_lo[3],
                         const int crse_region_hi[3], void *stream);

/* ---- Batched path: a device-resident work list walked by one persistent
 * kernel.  Replaces the loop of FillPatchInterp -> interp calls over fine
 * patches (CarpetX/src/fillpatch.cxx:116-119,179-180,213-216). */
typedef struct cxp_worklist_s *cxp_worklist;

int cxp_worklist_create(cxp_worklist *out);
int cxp_worklist_destroy(cxp_worklist wl);
int cxp_worklist_clear(cxp_worklist wl);
/* Adds one interp() call; same arguments and checks as cxp_prolongate. */
int cxp_worklist_append(cxp_worklist wl, cxp_op op, const cxp_fab *crse,
                        int crse_comp, const cxp_fab *fine, int fine_comp,
                        int ncomp, const int region_lo[3],
                        const int region_hi[3]);
/* Adds the ghost shell of a fine box: the fine FAB's box minus interior
 * [int_lo, int_hi], as six slabs (z-low/high full, y-low/high, x-low/high). */
int cxp_worklist_append_shell(cxp_worklist wl, cxp_op op, const cxp_fab *crse,
                              int crse_comp, const cxp_fab *fine,
                              int fine_comp, int ncomp, const int int_lo[3],
                              const int int_hi[3]);
/* A refinement level: nfabs FABs, each covering its valid region plus
 * ghost points, valid[6 i .. 6 i + 5] = valid lo[3], hi[3] of FAB i (valid
 * regions do not overlap), and the level's domain, the index range of its
 * valid points, in the FABs' index type.  Points outside the domain are
 * physical-boundary ghost points. */
typedef struct cxp_level {
  int nfabs;
  const cxp_fab *fabs;
  const int *valid;
  int domain_lo[3];
  int domain_hi[3];
} cxp_level;
/* Adds the ghost fill of a whole fine level from the next coarser level:
 * FillPatch_ProlongateGhosts (CarpetX/src/fillpatch.cxx:42-138) without its
 * MPI copies.  Every point of every fine FAB that lies in no fine FAB's
 * valid region (and within the fine domain grown by the FABs' ghost width)
 * is prolongated; points in another FAB's valid region are left alone, as
 * the reference leaves them to FillBoundary.  Coarse values come from the
 * coarse valid regions only, and, for coarse points outside the coarse
 * domain, from the coarse FABs' ghost points, which must hold the boundary
 * values already (the reference applies the boundary conditions to its
 * gathered coarse patch, fillpatch.cxx:107).  The fine patches are the
 * box differences of each FAB's box and the valid regions, z-slabs first
 * as in cxp_worklist_append_shell.  A patch whose CoarseBox lies in one
 * coarse FAB's available points reads that FAB in place; the others read a
 * coarse patch that the list gathers on the device at the start of every
 * run (the reference's ParallelCopy into its coarse patch, :86-93).
 * CXP_ERR_COARSE if a needed coarse point is in no FAB's available points. */
int cxp_worklist_append_level(cxp_worklist wl, cxp_op op,
                              const cxp_level *crse, int crse_comp,
                              const cxp_level *fine, int fine_comp, int ncomp);
/* Symmetry and boundary conditions of a level
 * (CarpetX/src/boundaries.hxx, boundaries_impl.hxx:150-476), one per face:
 * type[face][d], face 0 the lower and 1 the upper face in direction d.
 * The reference's reflection symmetry is CXP_BC_REFLECTION; otherwise the
 * group's boundary condition applies. */
#define CXP_BC_NONE 0
#define CXP_BC_DIRICHLET 1
#define CXP_BC_LINEAR_EXTRAPOLATION 2
#define CXP_BC_NEUMANN 3
#define CXP_BC_ROBIN 4
#define CXP_BC_REFLECTION 5
/* Periodic direction (both faces): coarse points outside the domain are the
 * periodic images of valid points, as the reference's coarse-patch copy
 * uses the coarse geometry's periodicity (fillpatch.cxx:91-93,175), with
 * period hi - lo + 1 (cell-centred) or hi - lo (vertex-centred) in the
 * domain's index range; a new level's fine domain is grown by the ghost
 * width in such directions (:160-170).  Fine valid boxes cover their
 * periodic images too: a ghost fill leaves points under another box's
 * image alone, and a remake copies old data from images (:222-223).  A
 * level's own periodic ghosts are left alone (the reference fills them by
 * FillBoundary). */
#define CXP_BC_PERIODIC 6
typedef struct cxp_bc {
  int type[2][3];
  const double *dirichlet_values; /* host, one per component, or null */
  const double *robin_values;     /* host, one per component, or null */
  const int *parities;            /* host, 3 per component (+1 / -1), or null */
  double xmin[3], dx[3];          /* coordinate of index 0 and spacing, as the
                                     reference's BoundaryCondition (robin) */
} cxp_bc;
/* cxp_worklist_append_level with the coarse level's boundary conditions
 * applied to the coarse patches, as FillPatch_ProlongateGhosts applies
 * them to its coarse patch (fillpatch.cxx:107) before interpolating.  Every
 * patch whose CoarseBox reaches outside the coarse domain reads a gathered
 * coarse patch: valid coarse data inside the domain, and outside it the
 * boundary values BoundaryCondition::apply computes from the patch's own
 * interior points (all directions at once, as on faces, edges and corners
 * there).  A face that such a patch reaches must have a type other than
 * CXP_BC_NONE, and the per-component arrays its type uses must be given
 * (CXP_ERR_ARG); at most 16 components.  bc = null is
 * cxp_worklist_append_level. */
int cxp_worklist_append_level_bc(cxp_worklist wl, cxp_op op,
                                 const cxp_level *crse, int crse_comp,
                                 const cxp_level *fine, int fine_comp,
                                 int ncomp, const cxp_bc *bc);
/* A new fine level's fill, FillPatch_NewLevel (CarpetX/src/fillpatch.cxx:
 * 140-181) without its MPI copies: each fine FAB's whole box within the
 * fine domain (valid and ghost points; the domain is not grown, as the
 * reference grows it in periodic directions only) is prolongated from the
 * coarse level, coarse patches as in cxp_worklist_append_level_bc (bc may
 * be null: the coarse FABs' boundary ghosts are read).  The fine level's
 * own boundary conditions are cxp_worklist_append_boundaries. */
int cxp_worklist_append_new_level(cxp_worklist wl, cxp_op op,
                                  const cxp_level *crse, int crse_comp,
                                  const cxp_level *fine, int fine_comp,
                                  int ncomp, const cxp_bc *bc);
/* A regridded fine level's fill, FillPatch_RemakeLevel (fillpatch.cxx:
 * 183-227): in each new fine FAB, the points within the fine domain grown
 * by the ghost width that lie in no valid region of the old fine level are
 * prolongated (coarse patches as in cxp_worklist_append_level_bc), and the
 * old level's valid data is copied into the rest of the FAB's box, on the
 * device at every run.  Both levels hold components fine_comp ..
 * fine_comp + ncomp - 1; the old and new FABs must not overlap in memory. */
int cxp_worklist_append_remake_level(cxp_worklist wl, cxp_op op,
                                     const cxp_level *crse, int crse_comp,
                                     const cxp_level *old_fine,
                                     const cxp_level *fine, int fine_comp,
                                     int ncomp, const cxp_bc *bc);
/* A level's own boundary conditions, in place, after the list's fills
 * (groupdata.apply_boundary_conditions(mfab), fillpatch.cxx:133,180,225;
 * BoundaryCondition::apply, boundaries_impl.hxx:150-437): in every FAB,
 * each point outside the domain in a direction whose face has a condition
 * takes its boundary value, computed from the FAB's own points.  Faces
 * with CXP_BC_NONE are left alone.  indextype: 1 for cell-centred
 * directions.  CXP_ERR_ARG if a FAB does not hold the sources of its
 * boundary points (the reference asserts this), or as for
 * cxp_worklist_append_level_bc. */
int cxp_worklist_append_boundaries(cxp_worklist wl, const cxp_level *level,
                                   const int indextype[3], int comp,
                                   int ncomp, const cxp_bc *bc);
/* The reference's `prolongate_per_group` parameter for the list's calls
 * (CarpetX/param.ccl:449-451; impl.hxx:2179-2187): 0 (the default) runs
 * ENO and fallback patches per variable, 1 with per-group decisions, as
 * cxp_prolongate's per_group.  Operators without branches give the same
 * values either way.  Takes effect at the next upload, which returns
 * CXP_ERR_NCOMP if a per-group patch has more than 16 components. */
int cxp_worklist_set_per_group(cxp_worklist wl, int per_group);
/* Builds tiles (using the tuned tile shapes) and copies the list to the
 * device; must precede cxp_prolongate_batched. */
int cxp_worklist_upload(cxp_worklist wl, void *stream);
int cxp_worklist_stats(cxp_worklist wl, int64_t *npatches, int64_t *ntiles,
                       int64_t *fine_points, int64_t *algo_bytes);
/* One launch of the persistent kernel over the whole list (plus one
 * point-kernel launch per ENO/fallback patch; first the gather of the level
 * fills' coarse patches and old-level copies, then the coarse patches'
 * boundary values; last the levels' own boundary conditions).  Launches of
 * one list never overlap: the list's dynamic-tail claim counters and
 * gathered patches serve one launch at a time, so a launch (of this, of
 * cxp_prolongate_per_patch or cxp_worklist_poison, or of a graph that
 * captured them) on another stream than the list's previous launch first
 * makes its stream wait for everything issued so far on that previous
 * stream (an event recorded there, then hipStreamWaitEvent), which must
 * therefore still exist.  Launches on one stream add no event. */
int cxp_prolongate_batched(cxp_worklist wl, void *stream);
/* The plain path over the same list: one cxp_prolongate launch per patch. */
int cxp_prolongate_per_patch(cxp_worklist wl, void *stream);
/* Writes `value` into every fine point of every patch of an uploaded list:
 * poison_invalid_gf (CarpetX/src/valid.cxx:157-215) for the regions a fill
 * will write, with CarpetX's FP64 poison pattern CXP_POISON_BITS
 * (valid.hxx:249-256) or any other value.  One store per point, no reads:
 * its run time is the floor of the list's write pattern. */
#define CXP_POISON_BITS 0xfff80000deadbeefULL
int cxp_worklist_poison(cxp_worklist wl, double value, void *stream);

/* ---- Reflux (CarpetX/src/schedule.cxx:2632-2716): the flux-register
 * correction of a coarse level where a finer level covers part of it, run
 * before restriction.  The reference calls AMReX's FluxRegister, made for
 * the fine level's cell-centred boxes with ratio 2 (driver.cxx:918-924):
 * CrseInit(coarse flux_d, d, ..., -1) and FineAdd(fine flux_d, d, ..., +1)
 * for d = 0, 1, 2, then Reflux(coarse state, 1.0, ..., coarse geometry)
 * (:2694-2704).  Here, for every register face F (a coarse face on the low
 * or high side of a coarsened fine valid box) and component:
 *   reg = crse_mult * Fc(F) + fine_mult * (each of the 4 fine faces of F);
 *   the coarse cell beyond F (index - 1 for a low face, the index for a high
 *   face, and its image across a periodic direction) gets
 *   -scale * reg / V (low) or +scale * reg / V (high), V = dx[0] dx[1] dx[2],
 *   if it is a valid cell of the coarse level (covered cells included, as in
 *   the reference; restriction overwrites them),
 * orientation by orientation in the order x-lo, y-lo, z-lo, x-hi, y-hi, z-hi.
 * The reference's multipliers are crse_mult = -1, fine_mult = 1, scale = 1;
 * AMReX's usual ones are -dt*A_coarse, dt*A_fine, 1.
 *   crse       the coarse level: cell-centred state FABs, valid boxes,
 *              domain (its extent is the period of periodic directions);
 *   crse_flux  per direction d, crse->nfabs FABs face-centred in d, each
 *              holding the faces of its state FAB's valid box;
 *   fine       the fine level: only its valid boxes are used (lo even, hi
 *              odd: coarsenable by 2);
 *   fine_flux  per direction d, fine->nfabs FABs holding the faces of their
 *              fine box in d.
 * Components comp .. comp + ncomp - 1 of every FAB.  Errors: CXP_ERR_ARG,
 * CXP_ERR_REGION (a FAB without its faces, cells or components),
 * CXP_ERR_COARSE (a register face no coarse flux FAB holds).  The plan holds
 * device copies of its pieces; cxp_reflux_run allocates nothing, so it can
 * be captured in a graph.  Destroy a plan after its launches completed. */
typedef struct cxp_reflux_s *cxp_reflux_plan;
int cxp_reflux_plan_create(const cxp_level *crse,
                           const cxp_fab *const crse_flux[3],
                           const cxp_level *fine,
                           const cxp_fab *const fine_flux[3], int comp,
                           int ncomp, const double dx[3],
                           const int periodic[3], double crse_mult,
                           double fine_mult, double scale,
                           cxp_reflux_plan *out);
int cxp_reflux_run(cxp_reflux_plan plan, void *stream);
/* Register faces of the plan, and the bytes one run moves (per face and
 * component: the coarse flux and 4 fine fluxes read, the target cell read
 * and written). */
int cxp_reflux_plan_stats(cxp_reflux_plan plan, int64_t *faces,
                          int64_t *bytes);
int cxp_reflux_plan_destroy(cxp_reflux_plan plan);

/* ---- hipGraph capture of a fill sequence.  A graph keeps the device
 * buffers and tiles of every list it launched.  Any append, clear,
 * cxp_worklist_set_per_group, upload, autotune or destroy of such a list
 * invalidates the graph: cxp_graph_launch then returns CXP_ERR_STATE and a
 * new capture is needed (after a regrid, say).  Lists launched inside a
 * capture begun by other means (torch.cuda.graph) are not tracked. */
typedef struct cxp_graph_s *cxp_graph;
int cxp_graph_capture_begin(void *stream);
int cxp_graph_capture_end(void *stream, cxp_graph *out);
int cxp_graph_launch(cxp_graph g, void *stream);
int cxp_graph_destroy(cxp_graph g);

/* ---- Tile-shape autotuner.  Tunes the list's separable size classes one
 * after another: every tile shape x z-chunk of a class is timed on the whole
 * list (the persistent kernel overlaps all classes), and a class keeps the
 * fastest only if it beats the best so far by more than 0.5%; passes repeat
 * (at most 3) until one changes nothing.  The table is
 * process-wide, keyed by (stencil triple, size class), and can be
 * saved/loaded as text.  Uploads the list again when done.  The timed
 * launches are the fill kernel alone (no gathers, copies or boundary
 * values) and write the list's fine points with values that mean nothing:
 * run the list after tuning. */
int cxp_autotune(cxp_worklist wl, int reps, void *stream);
int cxp_tune_save(const char *path);
int cxp_tune_load(const char *path);
/* Sets the table's entry for op's stencils and the size class of the
 * region [lo, hi]: tile shape cfg (0 .. cxp_tile_config_count() - 1) and
 * z-chunk zchunk (fine planes, >= 2; 0: the launch's default).  Applies to
 * per-patch launches and to lists uploaded afterwards.  Errors: CXP_ERR_ARG
 * for a null op or region, a shape out of range or a chunk of 1 or less
 * than 0.  No device work. */
int cxp_tune_set(cxp_op op, const int region_lo[3], const int region_hi[3],
                 int cfg, int zchunk);
/* Empties the table: every launch takes the default shapes again. */
int cxp_tune_clear(void);
int cxp_tile_config_count(void);
int cxp_tile_config(int cfg, int *tx, int *ty, int *zwindow);

#ifdef __cplusplus
}
#endif

#endif
