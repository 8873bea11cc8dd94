/*
 * flub.h -- C-ABI of the MI355X pressure-projection engine (libflub.so).
 *
 * The boundary replaces the Warp kernels and runtime services that the
 * reference's projection path launches from Python.  Every entry point takes
 * plain device pointers and sizes (no framework types), is stream-ordered on
 * the `stream` argument (NULL = the default stream), and returns an int:
 *   0 ok; < 0 argument/state error (FLUB_E* in csrc/common.h); > 0 hipError_t.
 * flub_last_error() returns the text of the last failure of the calling thread.
 *
 * Layout (identical to the reference's wp.array3d fields, const_param.py:114-143):
 *   cell fields  float32 [nx][ny][nz], k fastest
 *   u faces      float32 [nx+1][ny][nz]; v [nx][ny+1][nz]; w [nx][ny][nz+1]
 *   tags         int32   [nx][ny][nz], 0 AIR, 1 FLUID, 2 SOLID at the boundary
 *                (packed to one byte per cell inside the solver handle).
 * Grid sizes are limited to nx*ny*nz < 2^31 and faces < 2^31.
 */
#ifndef FLUB_H
#define FLUB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flub_mgpcg flub_mgpcg;

/* Smoother ids of MGPCG.solve / apply_preconditioner (multigrid/mgpcg.py:286-349):
 * "default" = tag-free _smooth/_restrict/compute_Ap with mean removal,
 * "phases"  = tagged _smooth_mcadams/_restrict_mcadams/compute_Ap_mcadams,
 * "fraction"= "phases" plus mean removal (zero_centered) as in solve(). */
enum { FLUB_SMOOTHER_DEFAULT = 0, FLUB_SMOOTHER_PHASES = 1, FLUB_SMOOTHER_FRACTION = 2 };

/* Storage precision of the V-cycle (the outer CG is always FP32 with FP64 dots).
 * FP32 is the reference V-cycle, bitwise.  FP16 / BF16 (no counterpart in the
 * reference) store level 0's r and pre-smoothed z in 16 bits, and the BF16
 * solve also level 1's when that level is colour-split; the 16-bit
 * pre-smoothing sweeps update in direct form, (r + sum of the neighbours) /
 * divd, with the hardware reciprocal; the outer loop is flexible CG. */
enum { FLUB_PREC_FP32 = 0, FLUB_PREC_FP16 = 1, FLUB_PREC_BF16 = 2 };

/* ---------------------------------------------------------------------------
 * Runtime services.  Replace Warp's array/runtime API the path relies on:
 * wp.empty/zeros/copy/from_numpy/.numpy()/.zero_()/.fill_() and
 * wp.ScopedTimer(synchronize=True) (const_param.py:114-150, water_3d.py:62).
 * ------------------------------------------------------------------------- */
const char *flub_last_error(void);
int flub_device_count(int *count);
int flub_malloc(void **ptr, size_t bytes);                       /* wp.empty */
int flub_free(void *ptr);
int flub_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);   /* wp.from_numpy */
int flub_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);   /* .numpy() */
int flub_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);   /* wp.copy */
int flub_memset(void *ptr, int byte, size_t bytes, void *stream);              /* .zero_() */
int flub_fill_f32(float *ptr, float value, size_t n, void *stream);           /* .fill_() */
int flub_fill_i32(int32_t *ptr, int32_t value, size_t n, void *stream);
int flub_stream_create(void **stream);
int flub_stream_destroy(void *stream);
int flub_stream_synchronize(void *stream);
int flub_device_synchronize(void);
int flub_event_create(void **event);
int flub_event_destroy(void *event);
int flub_event_record(void *event, void *stream);
int flub_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---------------------------------------------------------------------------
 * Cell classification (geometry.py:335-344 _tag_dirichlet, :390-397 _tag_neumann).
 * ------------------------------------------------------------------------- */
/* tag = phi > 0 ? AIR : FLUID over n cells. */
int flub_tag_dirichlet(const float *phi, int32_t *tag, size_t n, void *stream);
/* tag = SOLID where phi_rigid < 0, else unchanged. */
int flub_tag_neumann(const float *phi_rigid, int32_t *tag, size_t n, void *stream);

/* ---------------------------------------------------------------------------
 * Solid geometry: primitive signed distances, face fractions, rigid face
 * velocities (geometry.py:17-290,401-462; utils/util.py:168-235;
 * rigid_coupling.py:51-70).  Bodies are structure-of-arrays in device memory:
 * x[3n], q[4n] (quaternion x, y, z, w), gtype[n] (0 none, 1 quad, 2 circle,
 * 3 rect, 4 "x", 5 mesh: distance 2.0, resolved from cached level sets by the
 * caller), radius[3n], v[3n], omega[3n].  Arrays are limited to 2^32 points.
 * ------------------------------------------------------------------------- */
/* geometry.py:225-239 precompute_grid_point_phi_prims over an sx x sy x sz array:
 * phi at ((i + st_x) h, (j + st_y) h, (k + st_z) h) is the smallest body distance
 * below h (else 2.0, geometry.py:148-169); owner (may be NULL; the reference
 * writes it for staggered arrays) gets that body's index or -1. */
int flub_phi_prims(float *phi, int32_t *owner, int sx, int sy, int sz, float stx, float sty, float stz,
                   float h, int n_bodies, const float *x, const float *q, const int32_t *gtype,
                   const float *radius, void *stream);
/* geometry.py:264-290 phi_to_fraction for axis dim (1 u, 2 v, 3 w) of an nx x ny x nz
 * grid: phi holds the corner distances ((nx+1) x (ny+1) x (nz+1)), frac the face-centre
 * distances on entry and the face fractions (utils/util.py:168-228) on return. */
int flub_phi_to_fraction(const float *phi, float *frac, int nx, int ny, int nz, int dim, void *stream);
/* geometry.py:440-462 PrecomputeFaceFraction.precompute_fij_with_cache for primitive
 * bodies: face fractions and owners of all three axes, one pass per axis with the
 * corner and centre distances evaluated in registers (bitwise the staged sequence
 * phi_prims -> phi_to_fraction, without its (nx+1)(ny+1)(nz+1) corner array). */
int flub_face_fractions(float *fu, float *fv, float *fw, int32_t *owner_u, int32_t *owner_v, int32_t *owner_w,
                        int nx, int ny, int nz, float h, int n_bodies, const float *x, const float *q,
                        const int32_t *gtype, const float *radius, void *stream);
/* geometry.py:365-388 update_argmin, launched per cached level set by
 * PrecomputeFaceFraction.precompute_phi_with_cached_levelset (geometry.py:417-433):
 * the body's level set (lx x ly x lz values, spacing h_levelset) is sampled
 * tricubically (utils/util.py:98-104, 130-159) at p' = q^-1 (p - x) / h_levelset - ijk
 * for every point p = ((i + st_x) h, ...) of the sx x sy x sz array; where the sample
 * is below phi it replaces it, and owner (staggered arrays only; may be NULL) gets id.
 * x[3], q[4] (x, y, z, w), stagger[3] and ijk[3] are host arrays read during the call;
 * the level set holds at most 2^30 values. */
int flub_update_argmin(float *phi, int32_t *owner, int sx, int sy, int sz, const float *levelset, int lx, int ly,
                       int lz, int id, const float *x, const float *q, float h_levelset, const float *stagger,
                       const float *ijk, float h, void *stream);
/* rigid_coupling.py:51-70 precompute_solid_vel for axis d (1 u, 2 v, 3 w) over its
 * face array (sx, sy, sz): vel = (v + omega x (p - x))[d-1] of the owning body where
 * h0/100 < i*h0 < 1 - h0/100 and owner != -1, else 0. */
int flub_solid_vel(float *vel, const int32_t *owner, int sx, int sy, int sz, int d, float h0,
                   const float *x, const float *v, const float *omega, void *stream);

/* ---------------------------------------------------------------------------
 * Velocity extrapolation into the air, the step after the projection
 * (extrapolate.py:9-47; water_3d.py:70-73).  Marks: 1 = to extrapolate,
 * 0 = known, -1 = extrapolated in this pass (flushed back to 1).
 * ------------------------------------------------------------------------- */
/* utils/util.py:424-437 set_unknown over the nx x ny x nz cells: the face mark
 * arrays (u, v, w shapes) get phi >= 0 ? 1 : 0 at the cell index; their last
 * face layer is left as it is. */
int flub_set_unknown(const float *phi, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny, int nz, void *stream);
/* extrapolate.py:9-31 extrapolate_3d on one face array (sx, sy, sz) in place:
 * faces marked 1 take the mean of the faces marked 0 in their in-range
 * 27-neighbourhood and are marked -1 (0.0 and left at 1 when there are none). */
int flub_extrapolate_3d(float *q, int32_t *unknown, int sx, int sy, int sz, void *stream);
/* extrapolate.py:33-37 flush_1_3d: -1 -> 1 over n marks. */
int flub_flush_1_3d(int32_t *tag, size_t n, void *stream);
/* extrapolate.py:40-47 extrapolate_by_1 on the three face arrays of an
 * nx x ny x nz grid: each axis' extrapolate + flush as one kernel (marks -1 end
 * as 1, as after the reference's pair of launches). */
int flub_extrapolate_by_1(float *u, float *v, float *w, int32_t *tu, int32_t *tv, int32_t *tw, int nx, int ny,
                          int nz, void *stream);
/* water_3d.py:70-73 extrapolate: set_unknown from phi, then n_passes times
 * extrapolate_by_1 -- computed as min(n_passes, 2) passes, which gives the same
 * fields: after two passes the marks are back in {0, 1, owners >= 2} and every
 * later pass recomputes the same values. */
int flub_extrapolate(float *u, float *v, float *w, int32_t *tu, int32_t *tv, int32_t *tw, const float *phi, int nx,
                     int ny, int nz, int n_passes, void *stream);

/* ---------------------------------------------------------------------------
 * Advection, the step before the projection (advect.py:11-137): RK3 backtrace
 * through the MAC velocity (u, v, w of an nx x ny x nz grid) sampled
 * trilinearly, tricubic sampling of the advected field (utils/util.py:33-159).
 * Positions are in grid units; h is the cell size.
 * ------------------------------------------------------------------------- */
/* advect.py:58-112,133-137 advect_mc of one field q_in (gx x gy x gz, stagger st)
 * into q_out (not q_in): the reference's three launches (forward, backward,
 * correction with its min/max limiter) as one kernel, bitwise the same. */
int flub_advect_mc(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *q_in,
                   float *q_out, int gx, int gy, int gz, float stx, float sty, float stz, float dt, float h0,
                   void *stream);
/* advect.py:38-56 advect_phi: the tricubic value at the backtraced point when it
 * lies inside the grid, else h0. */
int flub_advect_phi(const float *u, const float *v, const float *w, int nx, int ny, int nz, const float *phi,
                    float *new_phi, float stx, float sty, float stz, float dt, float h0, void *stream);
/* advect.py:25-35 advect_simple_3d: trilinear value at the backtraced cell centre. */
int flub_advect_simple_3d(const float *u, const float *v, const float *w, const float *q, float *new_q, int nx,
                          int ny, int nz, float dt, float h, void *stream);

/* ---------------------------------------------------------------------------
 * Grid kernels of the projection (launched by example_fluid.py:84-94,120-147).
 * ------------------------------------------------------------------------- */
/* divergence.py:44-99 divergence_tag: div = (vr-vl)+(vt-vb)+(vn-vf), a face flux
 * replaced by the rigid velocity when the neighbour across it is SOLID and 0 on
 * the domain walls.  Walls are taken per axis (the reference's res=shape[0] is
 * the same on the cubic grids it supports). */
int flub_divergence_tag(const float *u, const float *v, const float *w, float *div,
                        const float *u_rigid, const float *v_rigid, const float *w_rigid,
                        const int32_t *tag, int nx, int ny, int nz, void *stream);
/* divergence.py:5-41 divergence_fractional_3d: the face fluxes weighted by the
 * open fractions fu/fv/fw plus the rigid fluxes weighted by (1 - fraction). */
int flub_divergence_fractional_3d(const float *u, const float *v, const float *w, float *div,
                                  const float *fu, const float *fv, const float *fw, const float *u_rigid,
                                  const float *v_rigid, const float *w_rigid, int nx, int ny, int nz,
                                  void *stream);
/* divergence.py:174-180 mask_with_tag: div = 0 where tag != FLUID. */
int flub_mask_with_tag(float *div, const int32_t *tag, size_t n, void *stream);
/* utils/util.py:162-165 scale_by_3d (and mgpcg.py:9-12 _scale_by_3d): q *= s. */
int flub_scale(float *q, float s, size_t n, void *stream);
/* subtract_gradient.py:38-68 subtract_gradient_tag for axis d (1=u, 2=v, 3=w):
 * uf/u_rigid have the face shape of that axis; p and tag are cell fields. */
int flub_subtract_gradient_tag(float *uf, const float *u_rigid, const float *p, const int32_t *tag,
                               int nx, int ny, int nz, int d, void *stream);
/* The same face update from an FP64 pressure p (a solver's FP64 solution x:
 * the reference's p0 / h), computed in FP64 and rounded once; no FP64 path in
 * the reference. */
int flub_subtract_gradient_tag_f64(float *uf, const float *u_rigid, const double *p, const int32_t *tag,
                                   int nx, int ny, int nz, int d, void *stream);
/* warp.utils.array_inner (mgpcg.py:5,391,...): *out_dev = sum a*b, accumulated
 * in FP64 by a deterministic two-stage reduction. */
int flub_inner(const float *a, const float *b, size_t n, double *out_dev, void *stream);
/* flub_inner without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * bitwise the same result; the FP64 partials live in a static device array of
 * 2048 doubles (16 KiB per GPU, held until the process ends) instead of memory
 * flub_inner allocates.  n is limited to FLUB_INNER_NOALLOC_MAX_N (larger n:
 * FLUB_ESHAPE).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (several streams or threads) race.  The stream must belong to the
 * current device. */
#define FLUB_INNER_NOALLOC_MAX_N 2147483647u
int flub_inner_noalloc(const float *a, const float *b, size_t n, double *out_dev, void *stream);

/* ---------------------------------------------------------------------------
 * MGPCG solver handle (multigrid/mgpcg.py:234-490).  The handle owns the level
 * pyramid (r, z, packed tags per level), x, p, Ap and the device scalars; no
 * allocation happens after create.  One handle per stream; not thread-safe.
 * ------------------------------------------------------------------------- */
/* MGPCG.__init__ (mgpcg.py:235-268): level l has shape (n >> l) per axis.
 * Requires every level extent >= 2 and nz >> (levels-1) even. */
int flub_mgpcg_create(flub_mgpcg **out, int nx, int ny, int nz, int n_mg_levels, int precision,
                      void *stream);
int flub_mgpcg_destroy(flub_mgpcg *h);
/* Attributes pre_and_post_smoothing / bottom_smoothing (mgpcg.py:238-239). */
int flub_mgpcg_set_smoothing(flub_mgpcg *h, int pre_and_post_smoothing, int bottom_smoothing);
/* No counterpart in the reference: the scale cs of the coarse-grid correction.
 * Every restriction stores cs times the sum of the eight children's residuals
 * (cs = 1, the default, is the reference's un-normalised sum, mgpcg.py:94-101,
 * and leaves every result bitwise unchanged); with cs < 1 each level's
 * prolongated correction shrinks by cs (cs = 0.25 is the Galerkin scale of
 * the piecewise-constant transfers).  Applies to every V-cycle of the handle
 * (FP32 and mixed, eager, graph and fused solves).  cs in (0, 4]. */
int flub_mgpcg_set_coarse_scale(flub_mgpcg *h, float cs);
/* No counterpart in the reference: level l of the V-cycle pre- and
 * post-smooths min(pre_and_post_smoothing << l, cap) red-black iterations
 * (cap 0, the default: the reference's doubling per level, mgpcg.py:301,341).
 * cap in [0, 1024]. */
int flub_mgpcg_set_sweep_cap(flub_mgpcg *h, int cap);
/* MGPCG.set_tags (mgpcg.py:360-369): tag (int32, may be NULL to keep level 0)
 * -> packed tags of every level (restrict_tag, mgpcg.py:162-174). */
int flub_mgpcg_set_tags(flub_mgpcg *h, const int32_t *tag);
/* MGPCG.init (mgpcg.py:276-284): zero z0, Ap, p, x; r0 = k * r. */
int flub_mgpcg_init(flub_mgpcg *h, const float *r, float k);
/* MGPCG.mask_interior (mgpcg.py:464-465,492-499). */
int flub_mgpcg_mask_interior(flub_mgpcg *h);
/* MGPCG.zero_centered (mgpcg.py:270-274,191-205). */
int flub_mgpcg_zero_centered(flub_mgpcg *h);
/* MGPCG.apply_preconditioner / graph_precondition (mgpcg.py:286-349); the
 * V-cycle is captured once per smoother into a HIP graph and replayed. */
int flub_mgpcg_apply_preconditioner(flub_mgpcg *h, int smoother);
int flub_mgpcg_graph_precondition(flub_mgpcg *h, int smoother);
/* compute_Ap_mcadams / compute_Ap (mgpcg.py:30-39,410-414): Ap = -A p. */
int flub_mgpcg_compute_Ap(flub_mgpcg *h, int smoother);
/* update_p/r/x (mgpcg.py:207-232,351-358) and update_alpha/beta (:212-222,372-376). */
int flub_mgpcg_update_p(flub_mgpcg *h);
int flub_mgpcg_update_r(flub_mgpcg *h);
int flub_mgpcg_update_x(flub_mgpcg *h);
int flub_mgpcg_update_alpha(flub_mgpcg *h, double eps);
int flub_mgpcg_update_beta(flub_mgpcg *h, double eps);
/* array_inner of two handle buffers into a handle scalar.
 * Buffers: 0 x, 1 p, 2 Ap, 3 r[0], 4 z[0]; scalars: 0 alpha, 1 beta, 2 sum, 3 old_zTr. */
int flub_mgpcg_inner(flub_mgpcg *h, int buf_a, int buf_b, int scalar_out);
/* MGPCG.solve (mgpcg.py:378-462).  max_iters = -1 for no limit; stops when
 * r.r < max(abs_tol, rel_tol * r0.r0) (squared norms).  *iters receives the
 * completed-iteration count, *rTr the last r.r.  verbose prints the residual
 * of every iteration as the reference does.  The colour-split solves ("phases"
 * on grids with nz % 8 == 0 and even nx, ny; every precision) accumulate x in
 * FP64 and, once the recursive r.r passes the test, recompute r = b - A x in
 * FP64: they stop when that true residual passes the same test (else r is
 * replaced and the solve goes on, at most 3 times), and *rTr is its squared
 * norm.  The other solves are the reference's FP32 algorithm. */
int flub_mgpcg_solve(flub_mgpcg *h, int max_iters, double eps, double abs_tol, double rel_tol,
                     int smoother, int verbose, int *iters, double *rTr);
/* compute_Ap_mcadams(Ap, p, tag) (mgpcg.py:36-39) on caller buffers of the
 * level-0 shape with the handle's tags: y = neighbor_dot_tag(x) (= -A x). */
int flub_mgpcg_apply_A(flub_mgpcg *h, const float *x, float *y);
/* Accuracy check, no counterpart in the reference (MGPCG.solve reports only
 * the recursive residual, mgpcg.py:393): *rr = ||b - neighbor_dot_tag(x)||^2
 * over the FLUID cells, the stencil evaluated in FP64 from the FP32 x and b
 * (caller buffers of the level-0 shape).  Synchronizes the handle's stream. */
int flub_mgpcg_residual_norm(flub_mgpcg *h, const float *x, const float *b, double *rr);
/* The same check for an FP64 solution (caller buffers of the level-0 shape,
 * x FP64): *rr = ||b - neighbor_dot_tag(x)||^2 over the FLUID cells, the
 * stencil in FP64; r (may be NULL) receives b - neighbor_dot_tag(x) rounded to
 * FP32.  Synchronizes the handle's stream.  No counterpart in the reference. */
int flub_mgpcg_true_residual_f64(flub_mgpcg *h, const double *x, const float *b, float *r, double *rr);
/* The last solve's x in FP64 (reference layout, caller buffer of the level-0
 * shape).  The colour-split solves (the "phases" smoother on grids with nz % 8
 * == 0 and even nx, ny) accumulate x in FP64 and hand back x rounded to FP32
 * in the x buffer; this returns the FP64 values until the next init, solve,
 * update_x or profile call.  Otherwise it returns the FP32 x widened.  No
 * counterpart in the reference (its x is FP32, mgpcg.py:207-210). */
int flub_mgpcg_solution_f64(flub_mgpcg *h, double *x);
/* One preconditioner application on caller buffers: r[0] <- r, V-cycle
 * (apply_preconditioner, mgpcg.py:296-349), z <- z[0]. */
int flub_mgpcg_precondition_into(flub_mgpcg *h, const float *r, float *z, int smoother);
/* Device pointers of the handle's buffers (x, p, Ap, r[l], z[l], tag[l]). */
int flub_mgpcg_buffer(flub_mgpcg *h, int which, int level, void **ptr);
/* divd[l] of the reference (mgpcg.py:51-58), computed from the packed tags into
 * a caller float buffer of the level's shape. */
int flub_mgpcg_compute_divd(flub_mgpcg *h, int level, float *divd);
/* The device scalars alpha, beta, sum, old_zTr, rTr (FP64), copied to host. */
int flub_mgpcg_scalars(flub_mgpcg *h, double out[5]);
int flub_mgpcg_level_shape(flub_mgpcg *h, int level, int shape[3]);
/* Measurement hook (the reference's wp.ScopedTimer, water_3d.py:62, at kernel
 * granularity): replays one kernel of the solve `reps` times on the handle's
 * stream between two HIP events and returns the mean ms per launch and the
 * kernel's algorithmic HBM bytes per launch.  Kernel ids (FLUB_PROF_*):
 * 0 level-0 RB half-sweep, 1 Ap + p.Ap, 2 r/x update + r.r, 3 whole V-cycle,
 * 4 restriction 0->1, 5 prolongation 1->0, 6 p update, 7 one temporally
 * blocked launch of four level-0 half-sweeps, 8/9 the half-sweep and
 * restriction of the colour-split level 0 the FP32 solve runs on, 10 its p
 * step (p = beta p + z, then Ap and the partials of p.Ap: one pass where the
 * grid allows, else two; 14 the same followed by the finish that sums the
 * partials into the `sum` scalar), 11 its r step (r -= alpha Ap
 * and r.r) (8-10: FLUB_ESTATE when the grid has no split level 0), 12 the
 * level-0 post-smoothing half-sweep of the mixed-precision split solve (FP32 z,
 * FP16/BF16 r; FLUB_ESTATE on an FP32 handle).  Leaves the handle's
 * vectors in an unspecified state (call before init / after reading x). */
int flub_mgpcg_profile(flub_mgpcg *h, int kernel, int reps, float *avg_ms, double *bytes);
/* In-solve timing (the same hook, inside the solve): arms HIP events around
 * the launches of kernel 8 (the colour-split level-0 plain half-sweep) that
 * the following split FP32 solves make, up to max_launches (<= 512) of them;
 * flub_mgpcg_probe_read returns their mean and minimum event-to-event time
 * (the kernel plus its dispatch) and the count, and disarms.  FLUB_ESTATE when
 * the handle has no colour-split FP32 level 0 or is already armed. */
int flub_mgpcg_probe(flub_mgpcg *h, int kernel, int max_launches);
int flub_mgpcg_probe_read(flub_mgpcg *h, float *avg_ms, float *min_ms, int *count);

/* ---------------------------------------------------------------------------
 * Whole projection (example_fluid.py:120-147 with water_3d.py:58-68 call_mgpcg):
 * divergence_tag + mask + scale(-1/h), set_tags, init(div, -h), mask_interior,
 * solve, p = x*h, subtract_gradient_tag d=1,2,3 with p/h, p *= h.
 * u, v, w are updated in place; p (cell field) receives the stored pressure
 * x*h; div (cell field, may be NULL) receives the scaled divergence.
 * ------------------------------------------------------------------------- */
int flub_project(flub_mgpcg *h, float *u, float *v, float *w, const float *u_rigid,
                 const float *v_rigid, const float *w_rigid, const int32_t *tag, float h_cell,
                 int max_iters, double abs_tol, double rel_tol, int smoother, float *p, float *div,
                 int *iters, double *rTr);

/* ---------------------------------------------------------------------------
 * Learned preconditioner kernels (no counterpart in the reference).  BF16
 * channels-last tensors x[n][d][h][w][c]; weights wt[cout][27*cin] (taps
 * (dz,dy,dx) row-major, cin fastest); FP32 accumulation on the matrix cores.
 * Supported (cin, cout): (16,16) (16,32) (32,32) (32,64) (64,64) (96,32)
 * (48,16) and the swapped pairs used by the data gradient.
 * ------------------------------------------------------------------------- */
/* 3x3x3 "same" convolution, optional bias (FP32, may be NULL) and ReLU; writes
 * BF16 y, or FP32 y32 when y32 is not NULL.  The data gradient is this call
 * on dY with the taps flipped and the channel roles swapped. */
int flub_conv3d_bf16(const void *x, const void *wt, const float *bias, void *y, float *y32, int nb, int d, int h,
                     int w, int cin, int cout, int relu, void *stream);
/* Weight gradient dw[cout][27*cin] (FP32) = sum over voxels of dy (x) x,
 * reduced deterministically through the caller's FP32 workspace
 * (work_floats >= 4 * cout * 27 * cin). */
int flub_conv3d_wgrad_bf16(const void *x, const void *dy, float *dw, float *work, size_t work_floats, int nb, int d,
                           int h, int w, int cin, int cout, void *stream);
/* The U-Net decoder's first convolution of a level over cat(up2(xa), xb) without
 * materialising it: channels [0, ca) are xa (nb x d/2 x h/2 x w/2 x ca, nearest 2x
 * upsampling), channels [ca, cin) are xb (nb x d x h x w x (cin-ca)).  Supported:
 * (cin, ca, cout) = (48, 32, 16), (96, 64, 32); w % 16 == 0, h, d multiples of 8. */
int flub_conv3d_upcat_bf16(const void *xa, int ca, const void *xb, const void *wt, const float *bias, void *y,
                           float *y32, int nb, int d, int h, int w, int cin, int cout, int relu, void *stream);
/* 2x2x2 average pooling, channels-last BF16 (c a multiple of 8; even sizes). */
int flub_pool2_bf16(const void *x, void *y, int nb, int d, int h, int w, int c, void *stream);
/* The network input (16 BF16 channels per voxel): r/s on FLUID voxels, the FLUID
 * indicator, zeros. */
int flub_pack_input_bf16(const float *r, const uint8_t *fluid, float s, void *x, size_t nvox, void *stream);
/* 1x1 output head with one channel: y[v] = b + sum_c x[v][c] w[c] (FP32 out). */
int flub_head_fwd(const void *x, const float *w, float b, float *y, size_t nvox, int c, void *stream);
/* The head followed by the preconditioner's output scaling: y[v] = mask[v] ?
 * (b + sum_c x[v][c] w[c]) * scale : 0. */
int flub_head_fwd_masked(const void *x, const float *w, float b, float *y, size_t nvox, int c, const uint8_t *mask,
                         float scale, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLUB_H */
