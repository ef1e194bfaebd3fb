/*
 * enlightning_hip.h -- C ABI of libenlightning_hip.so, the MI355X (gfx950)
 * implementation of Enlightning's per-patch update path.
 *
 * 1. Drop-in kernels.  Same symbols, argument order and by-reference
 *    argument passing as the reference's Fortran subroutines declared in
 *    enlightningfort.h:22-101 (SAMRAI_F77_FUNC(name,NAME) mangles to name_
 *    with gfortran/flang), so enlightning::singleStep (enlightning.C:507-578)
 *    and enlightning::tagGradientDetectorCells (enlightning.C:615-626) call
 *    them unchanged.
 *
 *    Memory: every array argument may be a device pointer (hipMalloc'd) or
 *    a host pointer.  Device arrays are used in place and the work is queued
 *    on the library stream (enl_set_stream; default: the null stream) --
 *    results are visible in stream order; call enl_sync() before reading
 *    them from the host.  Host arrays are staged through a device mirror and
 *    copied back before the call returns (the reference's synchronous
 *    contract; PCIe-bound).  Scalars and the small xlo/dx/tag_tolerance
 *    arrays are always host memory, as in the reference.
 *
 *    Errors: the reference subroutines have no error path.  These keep the
 *    void signature and report through enl_last_error(); with the
 *    environment variable ENL_ABORT_ON_ERROR=1 they abort instead
 *    (TBOX_ERROR semantics).
 *
 * 2. Patch-loop entry points (enl_*): the host-side shape of the reference's
 *    patch strategy around the kernels (physical boundary fill, the
 *    method-of-lines RK step), operating on device-resident patch data.
 */
#ifndef ENLIGHTNING_HIP_H
#define ENLIGHTNING_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rhs variants: one per fortran/rhs_*.m4 file (makefile:6 picks one per
 * reference build). */
enum {
  ENL_RHS_HYBRID_HIGH_LLF = 0, /* fortran/rhs_hybrid_high_llf.m4 (default) */
  ENL_RHS_HYBRID_HIGH = 1,     /* fortran/rhs_hybrid_high.m4 */
  ENL_RHS_HYBRID_LOW_LLF = 2,  /* fortran/rhs_hybrid_low_llf.m4 */
  ENL_RHS_HYBRID_LOW = 3,      /* fortran/rhs_hybrid_low.m4 */
  ENL_RHS_WENO_HIGH_LLF = 4,   /* fortran/rhs_weno_high_llf.m4 */
  ENL_RHS_WENO_HIGH = 5,       /* fortran/rhs_weno_high.m4 */
  ENL_RHS_WENO_LOW_LLF = 6,    /* fortran/rhs_weno_low_llf.m4 */
  ENL_RHS_WENO_LOW = 7,        /* fortran/rhs_weno_low.m4 */
  ENL_RHS_DRP = 8              /* fortran/rhs_drp.m4 */
};

/* ---- 1. drop-in kernels ------------------------------------------------ */

#define ENL_RHS_ARGS                                                                          \
  const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,               \
      const int *gcw0, const int *gcw1, const double *xlo, const double *dx,                  \
      const double *w_n, const double *source, double *rhs, const double *pres,               \
      const double *temp, const double *weno_alpha, const double *kappa, const double *mu,    \
      const double *muB, const double *R, const double *T_star_n, const double *T_star_o,     \
      const double *tau_n, const double *tau_o, const double *tag_tolerance, const int *level, \
      const double *c, const double *c_lr, const int *nequ

/* rhs: replaces SAMRAI_F77_FUNC(rhs,RHS), enlightningfort.h:23-49, called at
 * enlightning.C:507-533.  Writes all of rhs (nx*ny*nequ, interior).  w_n,
 * pres, temp are ghosted (gcw), source is interior.  rhs_ runs the variant
 * named by ENL_RHS_VARIANT (name such as "hybrid_high_llf" or its number),
 * default hybrid_high_llf (makefile:6). */
void rhs_(ENL_RHS_ARGS);
/* Per-variant aliases with the identical signature. */
void rhs_hybrid_high_llf_(ENL_RHS_ARGS); /* fortran/rhs_hybrid_high_llf.m4:28-51 */
void rhs_hybrid_high_(ENL_RHS_ARGS);     /* fortran/rhs_hybrid_high.m4:28-51 */
void rhs_hybrid_low_llf_(ENL_RHS_ARGS);  /* fortran/rhs_hybrid_low_llf.m4:28-51 */
void rhs_hybrid_low_(ENL_RHS_ARGS);      /* fortran/rhs_hybrid_low.m4:28-51 */
void rhs_weno_high_llf_(ENL_RHS_ARGS);   /* fortran/rhs_weno_high_llf.m4:27-50 */
void rhs_weno_high_(ENL_RHS_ARGS);       /* fortran/rhs_weno_high.m4:27-50 */
void rhs_weno_low_llf_(ENL_RHS_ARGS);    /* fortran/rhs_weno_low_llf.m4:27-50 */
void rhs_weno_low_(ENL_RHS_ARGS);        /* fortran/rhs_weno_low.m4:27-50 */
void rhs_drp_(ENL_RHS_ARGS);             /* fortran/rhs_drp.m4:26-49 */

/* rhs of any variant, plus an optional interior int map (nx*ny, may be
 * NULL) of the scheme each cell took: 0 = DRP, 1 = WENO
 * (rhs_hybrid_high_llf.m4:239-258).  Returns 0 or a negative error code. */
int enl_rhs_variant(int variant, ENL_RHS_ARGS, int *switch_map);

/* enl_rhs_variant without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * array must be device memory (host arrays would need staging buffers).
 * The level-0 hybrid variants' WENO list lives in a fixed per-GPU array of
 * 2^22 cells (32 MiB, held until the process ends); a patch of more interior
 * cells is rejected with an error.  Same kernels, same results as
 * enl_rhs_variant, bit for bit.  All calls on one GPU share its list: calls
 * that overlap on one GPU (other streams, other host threads) race. */
int enl_rhs_variant_noalloc(int variant, ENL_RHS_ARGS, int *switch_map);

/* runge_kutta: replaces SAMRAI_F77_FUNC(runge_kutta,RUNGE_KUTTA),
 * enlightningfort.h:51-64, called at enlightning.C:538-551.
 * w_n(interior) = alpha_1*w + alpha_2*w_n + beta*dt*rhs; ghosts untouched. */
void runge_kutta_(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                  const int *gcw0, const int *gcw1, const double *dt, const double *alpha_1,
                  const double *alpha_2, const double *beta, double *w_n, const double *w,
                  const double *rhs, const int *nequ);

/* update_state: replaces SAMRAI_F77_FUNC(update_state,UPDATE_STATE),
 * enlightningfort.h:66-88, called at enlightning.C:556-578.  Note the
 * argument order starts with dx, xlo.  Writes the interior of pres, temp. */
void update_state_(const double *dx, const double *xlo, const int *ifirst0, const int *ilast0,
                   const int *ifirst1, const int *ilast1, const int *gcw0, const int *gcw1,
                   const double *w_n, double *pres, double *temp, const double *rho0, const double *c,
                   const double *gamma1, const double *c_p, const double *p0, const double *T0,
                   const double *c_lr, const double *c_v, const double *R_tilde, const double *M,
                   const double *R, const int *nequ);

/* tag_cells: replaces SAMRAI_F77_FUNC(tag_cells,TAG_CELLS),
 * enlightningfort.h:90-101, called at enlightning.C:615-626.  Writes all of
 * tags (nx*ny int, interior); w_n and source are ghosted (gcw >= 1). */
void tag_cells_(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                const int *gcw0, const int *gcw1, int *tags, const double *w_n, const double *source,
                const int *refine_tag_val, const double *tag_tolerance, const int *nequ);

/* ---- library state ------------------------------------------------------ */

/* Stream for all subsequent launches (a hipStream_t; NULL = null stream). */
int enl_set_stream(void *hip_stream);
/* Waits for the library stream.  Returns 0 or a negative error code. */
int enl_sync(void);
/* Last error of this thread: returns its code (0 = none) and copies the
 * message into buf (NUL-terminated, truncated to n).  Clears it. */
int enl_last_error(char *buf, size_t n);
/* Library version string. */
const char *enl_version(void);

/* ---- 2. patch-loop entry points (device pointers only) ----------------- */

/* The Enlightning{constants} block (input.txt:40-59), the derived tau_n,
 * tau_o (enlightning.C:218-234) and tag_tolerance (input.txt:4). */
typedef struct enl_gas {
  double weno_alpha, c, gamma1, p0, T0, c_lr, rho0, R, c_p, T_star_n, T_star_o;
  double mu, muB, kappa, c_v, R_tilde, M, tau_n, tau_o;
  double tag_tolerance[3];
} enl_gas_t;

/* Physical boundary fill of one patch's ghost strips: replaces
 * enlightning::setPhysicalBoundaryConditions, enlightning.C:790-1013.
 * touch[4]: side (left, right, bottom, top) lies on the domain boundary;
 * reflect[4]: REFLECT (1) or ABSORB (0).  pres, temp, source may be NULL. */
int enl_fill_physical_bc(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw0, int gcw1,
                         const int *touch, const int *reflect, double *w_n, double *pres,
                         double *temp, double *source);

/* Absorbing envelopes of one patch: replaces
 * enlightning::setAbsorbingBoundaryConditions, enlightning.C:630-786.
 * absorb[4]: side (left, right, bottom, top) has bc_type ABSORB
 * (input.txt:14-17); width: bc_absorb_width (input.txt:18); patch_xlo: the
 * patch's lower corner, grid_xlo / grid_xhi: the domain's; dx: cell size.
 * Damps the interior of w_n, pres, temp (device arrays) toward rho0, 0,
 * T0*rho0, p0, T0 with the reference's factors and cell selection, sides in
 * the reference's order.  Returns 0 or a negative error code. */
int enl_absorb(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw0, int gcw1, const double *patch_xlo,
               const double *dx, const double *grid_xlo, const double *grid_xhi, const int *absorb, double width,
               const enl_gas_t *gas, double *w_n, double *pres, double *temp);

/* Source parameters of the reference input (input.txt:20-38): src_mode
 * (10 PULSE, 20 DRIVEN), src_type (10 POINT, 20 SINE, 30 LIGHTNING) as in
 * enlightning.C:55-61; for LIGHTNING, nseg segments of 4 coordinates
 * (x1 y1 x2 y2, as enl_read_lightning returns them; host memory). */
typedef struct enl_source {
  int mode, type;
  double amp, spread, pos[2], sine_amp, sine_freq, omega;
  int nseg;
  const double *coords;
} enl_source_t;

/* Mass source of one patch: replaces enlightning::addSource for a patch,
 * enlightning.C:2140-2307.  source: device array of the interior cells
 * (the interior context: cell (i,j) at source[(i-ifirst0) + sy*(j-ifirst1)],
 * sy = nx there).  POINT and SINE overwrite it, LIGHTNING merges each
 * segment into it (v + t - v/amp*t), DRIVEN POINT multiplies by
 * sin(omega*(iteration + 1 - pulse_time0)*dt).  Returns 0 or a negative
 * error code. */
int enl_add_source(int ifirst0, int ilast0, int ifirst1, int ilast1, const double *patch_xlo, const double *dx,
                   const enl_source_t *src, int iteration, int pulse_time0, double dt, double *source, long sy);

/* Lightning-channel file: replaces enlightning::readLightningInput,
 * enlightning.C:2333-2389.  Fills coords (4 per segment, up to max_segments;
 * NULL to count only) and returns the number of segments, or a negative
 * error code. */
int enl_read_lightning(const char *path, double *coords, int max_segments);

/* Microphones: the pressure read of enlightning::recordAudio,
 * enlightning.C:2392-2506, for one patch.  For each microphone k inside
 * [patch_xlo, patch_xhi] (bounds inclusive) out[k] = pres at cell
 * ((int)((x-xlo0)/dx0), (int)((y-xlo1)/dx1)) of the interior pressure array
 * (row stride sy) and hit[k] = 1; else hit[k] = 0.  The reference reads one
 * cell past the patch for a microphone on its upper bound; this reads the
 * last cell.  mic_x, mic_y, pres, out, hit: device arrays.  The per-step
 * files stay with the caller. */
int enl_sample_mics(int ifirst0, int ilast0, int ifirst1, int ilast1, const double *patch_xlo,
                    const double *patch_xhi, const double *dx, int nmics, const double *mic_x,
                    const double *mic_y, const double *pres, long sy, double *out, int *hit);

/* nsteps method-of-lines RK steps of one patch that covers the whole domain
 * (all four sides physical): per step ww <- w_n(interior), then per stage
 * the physical fill, rhs, runge_kutta and update_state, i.e. the per-patch
 * sequence of enlightning::singleStep (enlightning.C:459-583) inside the
 * SAMRAI MethodOfLinesIntegrator stage loop.  alpha_1/alpha_2/beta hold
 * `order` coefficients (input.txt:129-138).  ww and KK are device scratch of
 * nx*ny*5 doubles.  absorb_width > 0: after each stage's update_state the
 * absorbing envelopes of enl_absorb on the sides whose reflect flag is 0
 * (enlightning.C:581-582), the domain being [xlo, xlo + n*dx]; <= 0: none.
 * Returns 0 or a negative error code. */
int enl_advance_patch(int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                      const double *xlo, const double *dx, const enl_gas_t *gas, const int *reflect,
                      int order, const double *alpha_1, const double *alpha_2, const double *beta,
                      double dt, int nsteps, int level, double *w_n, double *pres, double *temp,
                      double *source, double *ww, double *KK, double absorb_width);

/* Bytes of device scratch enl_advance_patch_fused needs for this patch:
 * two more ghosted states (w_n, pres, temp) and an interior rhs. */
size_t enl_fused_scratch_bytes(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw);

/* Same steps and same results as enl_advance_patch (bit for bit), computed
 * with one fused kernel per stage: each cell's rhs, its runge_kutta update
 * and its update_state are applied in one pass (rhs is never stored for DRP
 * cells), and the stages ping-pong between w_n and two scratch states, so
 * the current -> scratch copy disappears.  scratch: device memory of at
 * least enl_fused_scratch_bytes().  On return w_n, pres and temp hold the
 * state on the interior.  Without an absorbing envelope (absorb_width 0, or
 * every side REFLECT) each stage kernel also writes the physical boundary
 * ghosts of its output, so the physical boundary runs once per call rather
 * than before every stage, and the returned ghosts hold the boundary of the
 * returned state; with the envelope they are not refreshed. */
int enl_advance_patch_fused(int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                            const double *xlo, const double *dx, const enl_gas_t *gas, const int *reflect,
                            int order, const double *alpha_1, const double *alpha_2, const double *beta,
                            double dt, int nsteps, int level, double *w_n, double *pres, double *temp,
                            double *source, void *scratch, size_t scratch_bytes, double absorb_width);

/* ---- 3. batched hierarchy (device patch table) -------------------------- */

/* One patch of a hierarchy: level, box in the level's index space
 * (ifirst0, ilast0, ifirst1, ilast1) and geometry (SAMRAI's
 * CartesianPatchGeometry xlo, dx). */
typedef struct enl_patch_desc {
  int level;
  int box[4];
  double xlo[2];
  double dx[2];
} enl_patch_desc_t;

enum { ENL_BATCH_PLAIN = 0, ENL_BATCH_FUSED = 1, ENL_BATCH_GRAPH = 2, ENL_BATCH_FP32 = 3, ENL_BATCH_FP32_GRAPH = 4 };

/* Creates a batch: one device pool holding, per patch, three ghosted states
 * [w_n(5), pres, temp] (SAMRAI CellData layout, 7 components), a ghosted
 * source and interior scratch.  ratio: refinement ratio between levels. */
int enl_batch_create(int npatch, const enl_patch_desc_t *patches, int gcw, int ratio, void **handle);
/* Ghost-fill schedule of one level, as (patch, local index) pairs; local
 * index = (i-ifirst0+g) + Nx*(j-ifirst1+g):
 *   copy:    destination <- same-level source cell, all components;
 *   interp:  destination <- linear refine of the coarse source cell,
 *            c + fx*(c[+x]-c[-x])/2 + fy*(c[+y]-c[-y])/2;
 *   edge:    destination <- same-patch source, side 1..4 (left, right,
 *            bottom, top) negating the normal momentum at REFLECT walls
 *            (enlightning.C:836-944);
 *   corner:  same, applied after the edges (enlightning.C:948-1008). */
int enl_batch_set_fill(void *handle, int level, int ncopy, const int *c_dp, const int *c_dl, const int *c_sp,
                       const int *c_sl, int ninterp, const int *i_dp, const int *i_dl, const int *i_sp,
                       const int *i_sl, const double *i_fx, const double *i_fy, int nedge, const int *e_p,
                       const int *e_d, const int *e_s, const int *e_n, int ncorner, const int *k_p,
                       const int *k_d, const int *k_s, const int *k_n);
/* Fine-to-coarse average after each step: (coarse patch, coarse local,
 * fine patch, fine local of the lower-left fine cell), ratio^2 cells. */
int enl_batch_set_coarsen(void *handle, int n, const int *cp, const int *cl, const int *fp, const int *fl);
/* Tile shape of the stage kernels on every level: 32x8, 16x16, 16x8 or 8x8. */
int enl_batch_set_tiles(void *handle, int bx, int by);
/* Tile shape of the stage kernels on one level (same choices). */
int enl_batch_set_level_tiles(void *handle, int level, int bx, int by);
/* How the stage kernel of one level packs its tiles into workgroups:
 * ENL_PACK_SINGLE, one tile per workgroup; ENL_PACK_STRIDED, a persistent
 * grid of wpc waves per SIMD (one workgroup = one tile's lanes), each
 * workgroup taking every P-th tile of its XCD's range; ENL_PACK_CHUNKED,
 * the same grid, each workgroup one contiguous run of tiles of equal
 * estimated cost.  Results do not depend on it. */
enum { ENL_PACK_SINGLE = 0, ENL_PACK_STRIDED = 1, ENL_PACK_CHUNKED = 2 };
int enl_batch_set_level_packing(void *handle, int level, int mode, int wpc);
/* Device pointers of a patch's current state and ghosted source. */
int enl_batch_patch_state(void *handle, int pid, double **w_n, double **pres, double **temp, double **source);
/* Copies a patch's current state (any argument may be NULL) between host
 * arrays of its ghosted size and the device: to_device = 1 uploads. */
int enl_batch_copy_state(void *handle, int pid, int to_device, double *w_n, double *pres, double *temp,
                         double *source);
/* Fills the source ghosts (constant over a run) and its interior copy. */
int enl_batch_prepare(void *handle, const int *reflect);
/* Zeroes every patch's source (ghosts and interior copy), as
 * enlightning::setSourceToZero (enlightning.C:2310-2329) does after a step
 * that added a source (main.C:476-480).  A fill kernel on the library
 * stream (no memset node), so it may also be captured into a graph. */
int enl_batch_zero_source(void *handle);
/* Ghost fill of the current state of every level. */
int enl_batch_fill(void *handle, const int *reflect);
/* pres, temp of every patch interior from w_n (update_state.m4). */
int enl_batch_update_state(void *handle, const enl_gas_t *gas);
/* nsteps method-of-lines steps of the whole hierarchy (one dt for every
 * level, no subcycling, enlightning.C:446-447).  mode: ENL_BATCH_PLAIN = the
 * drop-in entry points patch by patch; ENL_BATCH_FUSED = batched fused
 * stages; ENL_BATCH_GRAPH = the same captured once into a hipGraph and
 * replayed.  All three give bit-identical results. */
int enl_batch_advance(void *handle, int mode, int variant, const enl_gas_t *gas, const int *reflect, int order,
                      const double *alpha_1, const double *alpha_2, const double *beta, double dt, int nsteps);
/* Absorbing envelopes for the batch (enl_absorb on every patch after each
 * stage, enlightning.C:581-582), on the sides whose reflect flag is 0 in
 * each enl_batch_advance: grid_xlo / grid_xhi the domain, width the
 * bc_absorb_width.  width <= 0 turns them off.  Not available in the
 * FP32-perturbation modes (they return an error when a side absorbs). */
int enl_batch_set_absorb(void *handle, const double *grid_xlo, const double *grid_xhi, double width);
/* Regrid transfer: fills the new hierarchy new_handle from the old one
 * (gridding_algorithm->regridAllFinerLevels, main.C:489-502, which fills a
 * new level's SOLN variables w_n, pressure, temperature and source,
 * enlightning.C:259-281, from the old level where it existed and by
 * LINEAR_REFINE from the coarser level elsewhere).  Lists:
 *   same:    (new patch, old patch) pairs with the same level and box: the
 *            whole interior is copied;
 *   copy:    (new patch, new local) <- (old patch, old local): every SOLN
 *            component of an old cell of the same level;
 *   interp:  (new patch, new local) <- (old patch, old local): the linear
 *            refinement of an old coarser-level cell (same expression as
 *            the ghost fill's interp).
 * The old hierarchy's ghosts are refilled first (its interiors are not
 * written); afterwards the new hierarchy's source and state ghosts are
 * filled, so it is ready for enl_batch_advance.  The FP64 state is what
 * moves: after steps in the FP32-perturbation modes call
 * enl_batch_fp32_store on the old batch first.  Items that address cells
 * outside their patches' ghosted boxes (or refinement stencils that leave
 * the coarse patch) and pairs whose level or box differ are rejected before
 * anything runs. */
int enl_batch_transfer(void *new_handle, void *old_handle, const int *reflect, int nsame, const int *s_new,
                       const int *s_old, int ncopy, const int *c_dp, const int *c_dl, const int *c_sp,
                       const int *c_sl, int ninterp, const int *i_dp, const int *i_dl, const int *i_sp,
                       const int *i_sl, const double *i_fx, const double *i_fy);
/* tag_cells of every patch of one level (enlightning::tagGradientDetectorCells,
 * enlightning.C:586-628, the tagging step of regridAllFinerLevels) after a
 * ghost fill of the current state: tags receives, patch after patch in pid
 * order, each patch's nx*ny interior tags row by row (host or device
 * memory; host memory is filled before the call returns). */
int enl_batch_tag_cells(void *handle, int level, const int *reflect, int refine_tag_val,
                        const double *tag_tolerance, int *tags);
/* The same tags written into a map of the level's index space: map[j*nmap +
 * i] for every cell (i, j) of a patch of the level, 0 elsewhere (one launch
 * over all patches; host or device memory).  The tag input of the next
 * regrid's clustering (enl_plan_build). */
int enl_batch_tag_map(void *handle, int level, const int *reflect, int refine_tag_val, const double *tag_tolerance,
                      int *map, int nmap);

/* The schedules above as rectangles of cells: ni x nj cells; cell (a, b)
 * has destination local dl + a + b*Nx(dp) and source
 *   kind 0 (copy):   sl + a + b*Nx(sp), patch sp;
 *   kind 1 (refine): coarse cell sl + pa/r + (pb/r)*Nx(sp) of patch sp,
 *                    pa = (phase & 255) + a, pb = (phase >> 8) + b, fine
 *                    centre offsets fx = ((pa % r) + 0.5)/r - 0.5, fy alike;
 *   kind 1+s (physical side s = 1..4, sp = dp): sl + b*Nx (left, right) or
 *                    sl + a (bottom, top), the side's negation.
 * A coarsen rectangle lists coarse cells: dl + a + b*Nx(dp) averages the
 * r x r fine block whose lower-left cell is sl + a*r + b*r*Nx(sp). */
typedef struct enl_rect {
  int dp, sp, dl, sl, ni, nj, kind, phase;
} enl_rect_t;
/* A level's ghost-fill schedule from rectangles: n1 copy / refine / edge
 * rectangles, then n2 corner rectangles (a second pass).  The device
 * expands them into the lists enl_batch_set_fill would hold (same items,
 * rectangle order).  Rectangles outside their patches' ghosted boxes,
 * refinements whose +-1 stencil leaves the coarse box and level mismatches
 * are rejected before anything runs. */
int enl_batch_set_fill_rects(void *handle, int level, int n1, int n2, const enl_rect_t *rects);
/* The fine-to-coarse schedule from rectangles of coarse cells. */
int enl_batch_set_coarsen_rects(void *handle, int n, const enl_rect_t *rects);
/* enl_batch_transfer with the copy and refine items as rectangles (dp, dl
 * in the new batch; sp, sl in the old one). */
int enl_batch_transfer_rects(void *new_handle, void *old_handle, const int *reflect, int nsame, const int *s_new,
                             const int *s_old, int nrect, const enl_rect_t *rects);
int enl_batch_destroy(void *handle);
/* The SOLN patch data of every patch in one buffer, patch after patch in
 * pid order: per patch the planes selected by `planes` (bit k < 5: w_n
 * component k; bit 5: pressure; bit 6: temperature; bit 7: source), in that
 * order, each over the interior (ghosted = 0, nx*ny values row by row: the
 * VisIt plot quantities registerModelVariables registers,
 * enlightning.C:295-378) or over the ghosted box (ghosted = 1: a
 * checkpoint's patch data).  One launch; `out` is host or device memory of
 * exactly `count` doubles (checked). */
int enl_batch_gather(void *handle, int planes, int ghosted, double *out, long count);
/* The inverse of enl_batch_gather over the ghosted boxes: every selected
 * plane of every patch from `in` (a resumed checkpoint's patch data). */
int enl_batch_scatter(void *handle, int planes, const double *in, long count);
/* Device memory that destroyed batches gave back stays cached for the next
 * batch of a similar size (a regrid replaces the hierarchy's pool every few
 * steps); this returns the current device's cached blocks to the runtime. */
int enl_batch_trim(void);

/* ---- regrid planning on the host (gridding_algorithm->regridAllFinerLevels,
 * main.C:489-502; the refine and coarsen schedules enlightning.C:259-288
 * registers) -------------------------------------------------------------
 * A plan holds a two-level hierarchy's boxes and its schedules as flat
 * integer (and double) lists, in the layout enl_batch_set_fill,
 * enl_batch_set_coarsen and enl_batch_transfer take.  Boxes are 5 ints per
 * patch in pid order: level, ifirst0, ilast0, ifirst1, ilast1. */
/* Level 0: the n0 x n0 base grid in tiles of tile0 (a last tile narrower
 * than g joins the one before it); level 1, when any of the n0*n0 coarse
 * tags (row by row, may be NULL) is set: the tags dilated by tag_buffer,
 * clustered per block of fine_max/ratio coarse cells into bounding boxes
 * (at least fine_min/ratio coarse cells a side inside the block, extended
 * to the domain edge when within one coarse cell of it), refined by ratio.
 * previous (a plan or NULL) lends its level-0 fill schedule when its base
 * boxes are the same. */
int enl_plan_build(int n0, int ratio, int g, int tile0, const int *tags, int tag_buffer, int fine_max, int fine_min,
                   const void *previous, void **plan);
/* A plan of given boxes (a restart's), with its schedules. */
int enl_plan_from_boxes(int n0, int ratio, int g, int npatch, const int *boxes, void **plan);
/* The regrid transfer from the old hierarchy (old_plan, or when that is
 * NULL the n_old boxes old_boxes) into the plan's: unchanged patches whole
 * ("transfer.same.new/old"), every other interior cell copied from the old
 * patch of its level ("transfer.copy.dp/dl/sp/sl") or refined from the old
 * coarser level ("transfer.interp.dp/dl/sp/sl/fx/fy"). */
int enl_plan_transfer(void *plan, void *old_plan, int n_old, const int *old_boxes);
/* The same transfer as rectangles ("transfer.rects", enl_rect_t rows of at
 * most 1024 cells) beside the "same" pairs, for enl_batch_transfer_rects. */
int enl_plan_transfer_rects(void *plan, void *old_plan, int n_old, const int *old_boxes);
/* A list of the plan: "boxes", "fill.rects" and "fill.rects_n" (n1, n2) of
 * one level, "coarsen.rects", "fill.copy.p/dst/sp/sl", "fill.interp.p/dst/
 * cp/cl/fx/fy", "fill.edge.p/dst/src/neg", "fill.corner.p/dst/src/neg" (of
 * one level), "coarsen.cp/cl/fp/fl", the transfer lists.  The per-cell fill
 * and coarsen lists are built on first request.  *data points into the plan
 * (valid until enl_plan_free), *count is its length; returns 1 for a list
 * of doubles, 0 for ints (or an absent, empty list). */
int enl_plan_get(const void *plan, const char *name, int level, const void **data, long *count);
int enl_plan_free(void *plan);

/* ---- FP32-perturbation form of the batched hierarchy (no reference
 * counterpart; the precision form of the hot path) -------------------------
 * The state is stored as FP32 deviations from an FP64 ambient that depends
 * on height only, with FP64 RK accumulation of the carried state.  The rhs is
 * rhs_hybrid_high_llf's, evaluated in FP32 on the perturbations.  Steps run
 * through enl_batch_advance with mode ENL_BATCH_FP32 or ENL_BATCH_FP32_GRAPH
 * (variant rhs_hybrid_high_llf, RK order >= 2). */
/* Ambient of one level: rho, rho*s, rho*T_n on level rows row_lo ..
 * row_lo+nrows-1 (covering every patch's ghosted rows); p and T follow from
 * update_state's equation of state. */
int enl_batch_set_ambient(void *handle, int level, int row_lo, int nrows, const double *rho, const double *rhos,
                          const double *rhoTn, const enl_gas_t *gas);
/* Sets up the FP32 form: computes the ambient correction (FP64 minus FP32
 * rhs of the ambient) and converts the current FP64 state into the FP32 form.
 * Uses FP64 state blocks 1 and 2 as scratch. */
int enl_batch_fp32_prepare(void *handle, int variant, const enl_gas_t *gas, const int *reflect);
/* Current FP64 state -> FP32 form (after prepare). */
int enl_batch_fp32_load(void *handle);
/* FP32 form -> the current FP64 state (ambient + perturbation). */
int enl_batch_fp32_store(void *handle);
/* The learned switch (enl_switch_create) in place of the density sensor on
 * level 0 of the FP32 form (rhs_hybrid_high_llf.m4:239-256, the level-0
 * patches of the loop at enlightning.C:459-583): before level 0's rhs the
 * prefilter (on the full density) and the network (on the differences of
 * the perturbations rho' and p', the deviation from the ambient) run over
 * every level-0 patch at once and write one decision map per patch (0 DRP,
 * 1 WENO, 2 the sensor decides), which the stage kernel reads.  A cell that
 * is a candidate only through the ambient's own profile (its perturbation
 * pairs are provably smooth) is left to the sensor without a network call.  every_stage != 0 decides before each RK stage;
 * every_stage == 0 decides at the first stage of a step and keeps the map
 * for the step's other stages.  totals (device, 3 values, or NULL) gains
 * per decision the candidates (cells not provably DRP), the network's WENO
 * cells and the cells left to the sensor.  sw == NULL restores the sensor.  The switch must
 * outlive its use by this batch.  With lo = 0, hi = 1 and every_stage set,
 * every decision is the sensor's.  The hierarchy's ghost width must be at
 * least 3 (the features of a 5 x 5 window read one cell past it). */
int enl_batch_fp32_set_switch(void *handle, void *sw, float lo, float hi, int every_stage,
                              unsigned long long *totals);
/* The decision map of level-0 patch pid as the switch last wrote it
 * (nx*ny bytes, row by row, host or device memory). */
int enl_batch_fp32_decisions(void *handle, int pid, unsigned char *out);

/* ---- learned WENO/DRP switch (no reference counterpart; stands in for the
 * hybrid variants' density sensor, rhs_hybrid_high_llf.m4:239-256) --------
 * A prefilter keeps the cells where the sensor provably says DRP; on the
 * other cells an MLP 100-64-32-1 (BF16 weights and inputs, FP32 accumulation)
 * over a 5x5 window of normalised rho and p differences gives p(WENO):
 * p > hi -> WENO, p < lo -> DRP, otherwise the sensor decides.  Parameters
 * are packed FP32: W1[64][100], b1[64], W2[32][64], b2[32], w3[32], b3. */
#define ENL_SWITCH_NPARAM 8577
int enl_switch_create(void **handle);
int enl_switch_destroy(void *handle);
int enl_switch_set_params(void *handle, const float *params); /* host array; resets the optimiser */
int enl_switch_get_params(void *handle, float *params);       /* host array */
/* Features (BF16 bits, n x 100) of listed interior cells (i + j*nx) of a
 * patch state; device arrays. */
int enl_switch_features(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw, const double *w_n,
                        const double *pres, const enl_gas_t *gas, const int *cells, int n, unsigned short *X);
/* p(WENO) of n feature rows (device arrays). */
int enl_switch_forward(void *handle, const unsigned short *X, int n, float *p);
/* Weighted mean binary cross-entropy of a batch (labels y, weights wt with
 * sum wsum; device arrays) and its gradient (device, ENL_SWITCH_NPARAM). */
int enl_switch_gradients(void *handle, const unsigned short *X, const float *y, const float *wt, int n, float wsum,
                         float *grad, float *loss);
/* One Adam step on a batch. */
int enl_switch_train_step(void *handle, const unsigned short *X, const float *y, const float *wt, int n, float wsum,
                          float lr, float *loss);
/* Decisions (0 DRP, 1 WENO, 2 sensor) of a patch state into decide (nx*ny,
 * device); counts (host, optional): candidates, WENO, undecided.  gcw >= 3:
 * the features of a cell's 5 x 5 window read one cell past it.  The switch
 * keeps a candidate list and a feature map sized for the largest patch
 * decided so far. */
int enl_switch_decide(void *handle, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw, const double *w_n,
                      const double *pres, const enl_gas_t *gas, float lo, float hi, unsigned char *decide,
                      unsigned *counts);
/* enl_advance_patch_fused on level 0 of a hybrid variant with the learned
 * switch deciding level 0's cells: before every RK stage (every_stage != 0)
 * or at the first stage of each step, the map kept for the step's other
 * stages (every_stage == 0); totals (device, 3 x u64, optional) accumulate
 * the decision counts.  gcw >= 3, as for enl_switch_decide. */
int enl_advance_patch_learned(void *sw, float lo, float hi, int every_stage, unsigned char *decide,
                              unsigned long long *totals,
                              int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                              const double *xlo, const double *dx, const enl_gas_t *gas, const int *reflect, int order,
                              const double *alpha_1, const double *alpha_2, const double *beta, double dt, int nsteps,
                              double *w_n, double *pres, double *temp, double *source, void *scratch,
                              size_t scratch_bytes, double absorb_width);

/* ---- measurement hooks (no reference counterpart) ---------------------- */

/* Kernel classes timed by the patch-loop entry points. */
enum {
  ENL_KERNEL_RHS = 0,
  ENL_KERNEL_RK = 1,
  ENL_KERNEL_STATE = 2,
  ENL_KERNEL_BC = 3,
  ENL_KERNEL_COPY = 4,
  ENL_KERNEL_STAGE = 5, /* fused rhs + runge_kutta + update_state */
  ENL_NUM_KERNEL_KINDS = 6
};

/* Starts bracketing every launch of enl_advance_patch with HIP events on the
 * launch stream (the timing equivalent of the reference's t_rhs, t_rk,
 * t_state timers, enlightning.C:177-197).  capacity = launches to time. */
int enl_profile_start(int capacity);
/* Stops timing, waits for the events and returns, per kernel class, the
 * summed device time in ms and the number of launches. */
int enl_profile_stop(double *ms, int *count, int nkinds);

#ifdef __cplusplus
}
#endif
#endif /* ENLIGHTNING_HIP_H */
