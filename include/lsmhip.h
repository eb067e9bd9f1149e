/*
 * lsmhip.h -- entry points of the HIP level-set library that have no
 * counterpart in the reference: library control, the fused WENO5 + RHS +
 * TVD-RK stage kernels, ghost-fill dispatch (including the periodic fill the
 * serial reference lacks) and FP64-accumulated diagnostics.
 *
 * Include this with the reference-named headers (lsm_*3d*.h) of the drop-in.
 */
#ifndef LSMHIP_H
#define LSMHIP_H

#include "lsmhip_config.h"

LSMHIP_BEGIN_DECLS

/* ---- error side channel (the reference's kernels return nothing and check nothing) ---- */
#define LSMHIP_OK 0
#define LSMHIP_ERR_BOX 1   /* a fill/include box is not inside an array's ghost box (nothing is done) */
#define LSMHIP_ERR_HIP 2   /* a HIP runtime call failed */
#define LSMHIP_ERR_ARG 3   /* invalid argument, or index lists too short for a band */
#define LSMHIP_ERR_INDEX 4 /* index-list points whose stencil left a ghost box were skipped */
#define LSMHIP_ERR_SIZE 5  /* an array exceeds 2^31 elements */
int lsmhip_get_last_error(void);
const char *lsmhip_get_last_error_string(void);
void lsmhip_clear_error(void);
/* Reads (and clears) the device flag of index-list kernels; needed after
 * calls made in stream-ordered mode.  Returns nonzero when points were skipped. */
int lsmhip_check_device_errors(void);

/* ---- execution ---- */
/* stream: a hipStream_t (NULL = the null stream).  synchronous != 0 (the
 * default) waits for every call to complete, as the reference's calls do;
 * 0 leaves calls stream-ordered (results needed on the host, such as CFL
 * steps and integrals, still synchronise). */
void lsmhip_set_stream(void *stream, int synchronous);
void *lsmhip_get_stream(void);
int lsmhip_synchronize(void);
/* sizeof(LSMLIB_REAL) of the loaded library: 8 (liblsmhip.so) or 4 (liblsmhip_f32.so). */
int lsmhip_real_size(void);

/* ---- ghost cells ---- */
/* Applies one fill to the face group of the reference's C helpers
 * (src/boundary_conditions/lsm_boundary_conditions.h:60-80): 0..5 one face,
 * 6 x faces, 7 y faces, 8 z faces, 9 all faces in the order x_lo, x_hi,
 * y_lo, y_hi, z_lo, z_hi.  kind: 0 linear, 1 signed linear, 2 copy
 * (= homogeneous Neumann), 3 periodic.  Returns 0 or -1. */
int lsmhip_fill_ghosts(LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), int bdry_location_idx,
                       int kind);
/* Periodic fill of one face (index 0..5): ghost layers copy the interior
 * layers one period away; faces applied x, y, z fill edges and corners too. */
void lsmhip_periodic_face_(LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), const int *bdry_location_idx);

/* ---- fused stages: one TVD-RK stage of phi_t + H = 0 in one kernel ----
 * All arrays are device arrays on the ghost box gb (phi_stage with its ghost
 * cells filled); the stage writes phi_out on fb only.  stage 1, 2, 3 are the
 * TVD-RK3 stages (u_next = u + dt L(u);  3/4 u_n + 1/4 (u_1 + dt L(u_1));
 * 1/3 u_n + 2/3 (u_2 + dt L(u_2))), 22 is TVD-RK2 stage 2.  phi_n (the step's
 * initial field) may be NULL for stage 1.  FP64: bitwise equal to the drop-in
 * chain named in each comment.  FP32: the same scheme in FP32 storage and
 * arithmetic, arranged for fewer instructions.  Returns 0 or -1. */

/* HJ_WENO5 -> ZERO_OUT + ADD_[CONST_]NORMAL_VEL_TERM -> RK stage; vn_field NULL: constant vn. */
int lsmhip_weno5_stage_normal(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                              LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), int stage, LSMLIB_REAL vn,
                              const LSMLIB_REAL *vn_field, LSMLIB_REAL dt, LSMLIB_REAL dx, LSMLIB_REAL dy,
                              LSMLIB_REAL dz);
/* UPWIND_HJ_WENO5 -> ZERO_OUT + ADD_ADVECTION_TERM -> RK stage; velocity arrays on gb. */
int lsmhip_weno5_stage_advect(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                              LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), int stage, const LSMLIB_REAL *vel_x,
                              const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z, LSMLIB_REAL dt, LSMLIB_REAL dx,
                              LSMLIB_REAL dy, LSMLIB_REAL dz);
/* As above with the divergence-free deformation velocity
 *   u = 2 A(x) B(y) B(z) c, v = -B(x) A(y) B(z) c, w = -B(x) B(y) A(z) c,
 *   A = sin^2(pi s), B = sin(2 pi s),
 * evaluated in the kernel from per-axis device tables indexed from the ghost
 * box origin: tabs = {A_x, A_y, A_z, B_x, B_y, B_z} (host array of 6 device
 * pointers); c is the time factor.  No velocity array is read. */
int lsmhip_weno5_stage_advect_deform(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                                     LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), int stage,
                                     const LSMLIB_REAL *const *tabs, LSMLIB_REAL c, LSMLIB_REAL dt, LSMLIB_REAL dx,
                                     LSMLIB_REAL dy, LSMLIB_REAL dz);
/* HJ_WENO5 -> COMPUTE_REINITIALIZATION_EQN_RHS -> RK stage. */
int lsmhip_weno5_stage_reinit(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                              LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), int stage, const LSMLIB_REAL *phi0,
                              int use_phi0_for_sgn, LSMLIB_REAL dt, LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz);
/* Writes the deformation velocity of lsmhip_weno5_stage_advect_deform into
 * three arrays on gb over box b. */
int lsmhip_deform_velocity_fill(LSMLIB_REAL *vel_x, LSMLIB_REAL *vel_y, LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(gb),
                                LSMHIP_BOX_ARGS(b), const LSMLIB_REAL *const *tabs, LSMLIB_REAL c);
/* k planes a fused-stage block marches through (default 64). */
void lsmhip_set_fused_kchunk(int planes);
/* FP32 library: 1 (default) selects the stage kernels that form a WENO5
 * window per cell, 2 the kernels that form each position's WENO5 terms once and
 * share them between neighbouring cells (the same scheme with different
 * rounding; fewer instructions but more registers, and slower on gfx950; kept
 * for comparison).  Returns the previous choice; ignored by the FP64 library. */
int lsmhip_set_fused_f32_kernels(int which);

/* ---- fused narrow-band stages: one TVD-RK stage of normal motion on a band ----
 * The band as the reference's local driver builds it (LSM3D_DETERMINE_NARROW_BAND
 * with at least three levels, boundary layers marked): index_x/y/z addressed
 * from element nlo_index0, level-0 entries nlo_index0..nhi_index0, narrow_band
 * on nb_gb, phi_stage / phi_n / vn_field / phi_out on gb.  One thread per
 * level-0 point with mark <= mark_fb; the differences of the WENO5 stencil
 * take the 'big' value where a neighbour's mark exceeds mark_D1, as
 * HJ_WENO5_LOCAL stores them.  A point whose stencil reaches a cell outside
 * the band (mark 0) is skipped and reported (LSMHIP_ERR_INDEX).  Returns 0 or -1. */

/* HJ_WENO5_LOCAL (lsm3dhjweno5local_, levels 0..3) -> ZERO_OUT_LOCAL ->
 * ADD_[CONST_]NORMAL_VEL_TERM_LOCAL -> MULTIPLY_CUT_OFF_LSE_RHS_LOCAL(phi_stage,
 * beta, gamma; none when gamma <= 0) -> TVD_RK3_STAGEn_LOCAL (stage 1, 2, 3; 22 =
 * TVD-RK2 stage 2).  FP64: bitwise equal to that chain; phi_n may be NULL for stage 1.
 * nhi_copy >= nlo_index0 folds in COPY_DATA_LOCAL(phi_out, phi_n) over entries
 * nlo_index0..nhi_copy (e.g. levels 0..3) before the stage: the entries the
 * stage does not update (beyond nhi_index0, or marked above mark_fb) take
 * phi_n (phi_stage at stage 1); nhi_copy < nlo_index0 copies nothing. */
int lsmhip_weno5_local_stage_normal(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                                    LSMHIP_BOX_ARGS(gb), int stage, LSMLIB_REAL vn, const LSMLIB_REAL *vn_field,
                                    LSMLIB_REAL dt, LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                    const int *index_x, const int *index_y, const int *index_z, int nlo_index0,
                                    int nhi_index0, int nhi_copy, const unsigned char *narrow_band,
                                    LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb, unsigned char mark_D1,
                                    LSMLIB_REAL beta, LSMLIB_REAL gamma);
/* HJ_WENO5_LOCAL -> COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL (vn_field on gb; only
 * points with |grad phi| > zero_tol count) or, with vn_field NULL,
 * COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL in its whole-grid form: dt = cfl /
 * (max over level-0 points of H/dx + tiny).  FP64: bitwise equal to that chain. */
int lsmhip_weno5_local_normal_dt(LSMLIB_REAL *dt, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMLIB_REAL vn,
                                 const LSMLIB_REAL *vn_field, LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                 LSMLIB_REAL cfl_number, const int *index_x, const int *index_y,
                                 const int *index_z, int nlo_index0, int nhi_index0,
                                 const unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb,
                                 unsigned char mark_D1);
/* lsmhip_weno5_local_normal_dt, and the RHS the first stage of the same chain
 * would form from phi (ZERO_OUT_LOCAL -> ADD_[CONST_]NORMAL_VEL_TERM_LOCAL ->
 * MULTIPLY_CUT_OFF_LSE_RHS_LOCAL(phi, beta, gamma; none when gamma <= 0)) at
 * each level-0 point, stored compactly: rhs[t] for entry nlo_index0 + t,
 * rhs holding nhi_index0 - nlo_index0 + 1 reals on the device.  Entries of
 * points the stage skips (mark above mark_fb, or outside the ghost boxes) are
 * left as they are; a point whose stencil reaches outside the band is reported
 * as in lsmhip_weno5_local_normal_dt and its entry marked to be skipped.  The
 * CFL pass already computes that point's derivatives, so the first stage need
 * not compute them again (lsmhip_local_stage_from_rhs). */
int lsmhip_weno5_local_normal_dt_rhs(LSMLIB_REAL *dt, LSMLIB_REAL *rhs, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb),
                                     LSMLIB_REAL vn, const LSMLIB_REAL *vn_field, LSMLIB_REAL dx, LSMLIB_REAL dy,
                                     LSMLIB_REAL dz, LSMLIB_REAL cfl_number, const int *index_x, const int *index_y,
                                     const int *index_z, int nlo_index0, int nhi_index0,
                                     const unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb),
                                     unsigned char mark_fb, unsigned char mark_D1, LSMLIB_REAL beta,
                                     LSMLIB_REAL gamma);
/* The stage of lsmhip_weno5_local_stage_normal (same stages, points, copies
 * over nlo_index0..nhi_copy and RK combination) with the RHS taken from rhs as
 * lsmhip_weno5_local_normal_dt_rhs stored it instead of computed: with that
 * rhs (of phi_stage) the result is bitwise the fused stage's, and so the
 * chain's in FP64.  An entry marked to be skipped is skipped and reported
 * (LSMHIP_ERR_INDEX).  Returns 0 or -1. */
int lsmhip_local_stage_from_rhs(LSMLIB_REAL *phi_out, const LSMLIB_REAL *phi_stage, const LSMLIB_REAL *phi_n,
                                LSMHIP_BOX_ARGS(gb), int stage, const LSMLIB_REAL *rhs, LSMLIB_REAL dt,
                                const int *index_x, const int *index_y, const int *index_z, int nlo_index0,
                                int nhi_index0, int nhi_copy, const unsigned char *narrow_band,
                                LSMHIP_BOX_ARGS(nb_gb), unsigned char mark_fb);

/* ---- reverse mode of the normal-motion stage (for gradient-based inverse problems) ----
 * Reverse of lsmhip_weno5_stage_normal for stage 1, 2, 3 (TVD-RK3) or 22
 * (TVD-RK2 stage 2): with out = a u_c + b (u_s + dt L(u_s)), L = -V sqrt(G),
 * given lam_out = dJ/d out on fb it adds b lam_out + (b dt dL/du_s)^T lam_out
 * to lam_us (ghost box), a lam_out to lam_uc (fb; unused for stage 1, may be
 * NULL) and -b dt sqrt(G) lam_out to lam_vn (fb; only with vn_field, may be
 * NULL).  scratch: 21 * |gb| reals.  Derivatives are those of the implemented
 * WENO5/Godunov functions (argmax of the WENO eps, the selected Godunov branch). */
int lsmhip_weno5_stage_normal_adjoint(LSMLIB_REAL *lam_us, LSMLIB_REAL *lam_uc, LSMLIB_REAL *lam_vn,
                                      const LSMLIB_REAL *lam_out, const LSMLIB_REAL *phi_stage, LSMHIP_BOX_ARGS(gb),
                                      LSMHIP_BOX_ARGS(fb), int stage, LSMLIB_REAL vn, const LSMLIB_REAL *vn_field,
                                      LSMLIB_REAL dt, LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                      LSMLIB_REAL *scratch);
/* Adjoint of the periodic ghost fill: adds each ghost entry to the interior
 * cell it copies, then zeroes the ghosts. */
int lsmhip_periodic_fill_adjoint(LSMLIB_REAL *lam, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb));
/* J = sum_fb (H_eps(phi) - H_eps(target))^2 dx dy dz in FP64, and dJ/dphi into
 * grad (gb, zero off fb) when grad is not NULL. */
int lsmhip_heaviside_mismatch(double *loss, LSMLIB_REAL *grad, const LSMLIB_REAL *phi, const LSMLIB_REAL *target,
                              LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), double dx, double dy, double dz,
                              double epsilon);

/* ---- diagnostics ---- */
/* Smoothed volume of {phi < 0} (the reference's lsm3dVolumeRegionPhiLessThanZero
 * integrand) accumulated and returned in FP64 whatever LSMLIB_REAL is. */
int lsmhip_volume_phi_lt0_f64(double *volume, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb), LSMHIP_BOX_ARGS(ib),
                              const double *dx, const double *dy, const double *dz, const double *epsilon);

/* ---- allocation-free variant ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): lsm3dvolumeregionphilessthanzero_ (the same kernels and block
 * count, so the same result bit for bit) with its per-block partial sums and
 * result in a file-scope device array of 64 KiB that each GPU holds until the
 * process ends, instead of library workspace.  It calls no hipMalloc or
 * hipFree.  Boxes of 2^31 cells or more are refused before anything is
 * launched (LSMHIP_ERR_SIZE).  Calls that overlap on one GPU (several host
 * threads or streams) race on that array.  Returns 0 or -1. */
int lsmhip_volume_region_phi_lt0_noalloc(LSMLIB_REAL *volume, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                         LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                         const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);

LSMHIP_END_DECLS

#endif
