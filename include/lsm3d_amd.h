/*
 * lsm3d_amd.h -- C ABI of the gfx950 (MI355X) level-set update library.
 *
 * Two builds share these declarations, as LSMLIB's --enable-float switch
 * does (configure.ac:258-263): liblsm3d_amd_double.so with LSMLIB_REAL =
 * double and liblsm3d_amd_float.so with LSMLIB_REAL = float.
 *
 * Part 1 -- drop-in toolbox entry points.  Each keeps the symbol name and the
 * by-reference Fortran-77 prototype of the LSMLIB routine it replaces (cited
 * per function), so the serial package, the SAMRAI patch loops and the MATLAB
 * MEX wrappers link against this library unchanged.  Arrays are column-major
 * over their own ghost box (i fastest).  Arrays may be device memory (the
 * kernels run in place) or host memory (staged through device buffers).  Calls
 * complete before returning unless lsm_amd_set_sync(0) is in effect.
 * Differences from the reference: the D1/D2/D3 scratch arguments are accepted
 * and not written; out-of-range boxes are reported instead of being undefined
 * behaviour; sums (volume integrals) accumulate in double precision.
 *
 * Part 2 -- runtime control (stream, synchronisation, error channel).
 * Part 3 -- device-resident fused kernels (this library's own API, no
 * counterpart in the reference): a whole TVD-RK stage of the advection or
 * reinitialization equation in one pass, device ghost fills, and a batched
 * patch-list advance.
 * Part 4 -- reverse mode of the fused advection stage (no counterpart in the
 * reference) and a smoothed-Heaviside mismatch loss, for gradient-based
 * callers.
 */
#ifndef LSM3D_AMD_H
#define LSM3D_AMD_H

#ifndef LSMLIB_REAL
#ifdef LSM_REAL
#define LSMLIB_REAL LSM_REAL
#else
#define LSMLIB_REAL double
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* the library is built with hidden visibility; everything declared here is
 * exported */
#pragma GCC visibility push(default)

/* six by-reference box bounds: ilo, ihi, jlo, jhi, klo, khi */
#define LSM_AMD_BOX(p)                                                   \
  const int *ilo_##p, const int *ihi_##p, const int *jlo_##p,           \
      const int *jhi_##p, const int *klo_##p, const int *khi_##p

/* ======================================================================
 * Part 1: drop-in LSMLIB toolbox routines
 * ====================================================================== */

/* --- spatial derivatives: src/toolbox/spatial_derivatives/lsm_spatial_derivatives3d.h --- */
/* LSM3D_HJ_ENO1, lsm_spatial_derivatives3d.h:83 */
void lsm3dhjeno1_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                  LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                  LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                  LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                  LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                  const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_HJ_ENO2, lsm_spatial_derivatives3d.h:159 */
void lsm3dhjeno2_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                  LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                  LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                  LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                  LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), LSMLIB_REAL *D2, LSM_AMD_BOX(D2_gb),
                  LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                  const LSMLIB_REAL *dz);
/* LSM3D_HJ_ENO3, lsm_spatial_derivatives3d.h:243 */
void lsm3dhjeno3_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                  LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                  LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                  LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                  LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), LSMLIB_REAL *D2, LSM_AMD_BOX(D2_gb),
                  LSMLIB_REAL *D3, LSM_AMD_BOX(D3_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                  const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_HJ_WENO5, lsm_spatial_derivatives3d.h:332 */
void lsm3dhjweno5_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                   LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                   LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                   LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                   LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                   const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_UPWIND_HJ_ENO1, lsm_spatial_derivatives3d.h:395 */
void lsm3dupwindhjeno1_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                        LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                        const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                        const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb), LSMLIB_REAL *D1,
                        LSM_AMD_BOX(D1_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                        const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_UPWIND_HJ_ENO2, lsm_spatial_derivatives3d.h:459 */
void lsm3dupwindhjeno2_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                        LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                        const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                        const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb), LSMLIB_REAL *D1,
                        LSM_AMD_BOX(D1_gb), LSMLIB_REAL *D2, LSM_AMD_BOX(D2_gb), LSM_AMD_BOX(fb),
                        const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_UPWIND_HJ_ENO3, lsm_spatial_derivatives3d.h:531 */
void lsm3dupwindhjeno3_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                        LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                        const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                        const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb), LSMLIB_REAL *D1,
                        LSM_AMD_BOX(D1_gb), LSMLIB_REAL *D2, LSM_AMD_BOX(D2_gb), LSMLIB_REAL *D3,
                        LSM_AMD_BOX(D3_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                        const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_UPWIND_HJ_WENO5, lsm_spatial_derivatives3d.h:608 */
void lsm3dupwindhjweno5_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                         LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                         const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                         const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb), LSMLIB_REAL *D1,
                         LSM_AMD_BOX(D1_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
                         const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* LSM3D_CENTRAL_GRAD_ORDER2, lsm_spatial_derivatives3d.h:673 (phi read at fb +- 1) */
void lsm3dcentralgradorder2_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                             LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                             LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                             const LSMLIB_REAL *dz);
/* LSM3D_CENTRAL_GRAD_ORDER4, lsm_spatial_derivatives3d.h:723 (phi read at fb +- 2) */
void lsm3dcentralgradorder4_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                             LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                             LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                             const LSMLIB_REAL *dz);
/* LSM3D_LAPLACIAN_ORDER2, lsm_spatial_derivatives3d.h:777 */
void lsm3dlaplacianorder2_(LSMLIB_REAL *laplacian_phi, LSM_AMD_BOX(laplacian_phi_gb),
                           const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(fb),
                           const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_GRADIENT_MAGNITUDE, lsm_spatial_derivatives3d.h:939 */
/* qualifiers as the reference header declares them */
void lsm3dgradientmagnitude_(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                             const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                             LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(fb));

/* LSM3D_PHI_UPWIND_GRAD_F, lsm_spatial_derivatives3d.h:831 (phi read at fb +- 1;
 * where phi == 0 the outputs are left unchanged, as in the reference) */
/* qualifiers as the reference header declares them */
void lsm3dphiupwindgradf_(LSMLIB_REAL *F_x, LSMLIB_REAL *F_y, LSMLIB_REAL *F_z, LSM_AMD_BOX(grad_F_gb),
                          LSMLIB_REAL *F_x_plus, LSMLIB_REAL *F_y_plus,
                          LSMLIB_REAL *F_z_plus, LSM_AMD_BOX(grad_F_plus_gb),
                          LSMLIB_REAL *F_x_minus, LSMLIB_REAL *F_y_minus,
                          LSMLIB_REAL *F_z_minus, LSM_AMD_BOX(grad_F_minus_gb),
                          const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(fb));
/* LSM3D_AVERAGE_GRAD_PHI, lsm_spatial_derivatives3d.h:893 */
/* qualifiers as the reference header declares them */
void lsm3daveragegradphi_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                          LSM_AMD_BOX(grad_phi_gb), LSMLIB_REAL *phi_x_plus,
                          LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                          LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                          LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                          LSM_AMD_BOX(grad_phi_minus_gb), LSM_AMD_BOX(fb));

/* --- level set equation RHS: src/toolbox/level_set_evolution/lsm_level_set_evolution3d.h --- */
/* LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS, lsm_level_set_evolution3d.h:64 */
void lsm3dzerooutlevelseteqnrhs_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb));
/* LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:94 */
void lsm3daddadvectiontermtolserhs_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                    const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
                                    const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                    const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb),
                                    LSM_AMD_BOX(fb));
/* LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:151 */
void lsm3daddnormalveltermtolserhs_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
                                    const LSMLIB_REAL *phi_z_plus,
                                    LSM_AMD_BOX(grad_phi_plus_gb),
                                    const LSMLIB_REAL *phi_x_minus,
                                    const LSMLIB_REAL *phi_y_minus,
                                    const LSMLIB_REAL *phi_z_minus,
                                    LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *vel_n,
                                    LSM_AMD_BOX(vel_gb), LSM_AMD_BOX(fb));
/* LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:217 */
void lsm3daddconstnormalveltermtolserhs_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                         const LSMLIB_REAL *phi_x_plus,
                                         const LSMLIB_REAL *phi_y_plus,
                                         const LSMLIB_REAL *phi_z_plus,
                                         LSM_AMD_BOX(grad_phi_plus_gb),
                                         const LSMLIB_REAL *phi_x_minus,
                                         const LSMLIB_REAL *phi_y_minus,
                                         const LSMLIB_REAL *phi_z_minus,
                                         LSM_AMD_BOX(grad_phi_minus_gb),
                                         const LSMLIB_REAL *vel_n, LSM_AMD_BOX(fb));
/* LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:277:
 * lse_rhs += b * kappa |grad phi| from central first derivatives (grad_phi_gb)
 * and second derivatives (grad2_phi_gb) */
/* qualifiers as the reference header declares them */
void lsm3daddconstcurvtermtolserhs_(const LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                    const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
                                    const LSMLIB_REAL *phi_xx, const LSMLIB_REAL *phi_xy,
                                    const LSMLIB_REAL *phi_xz, const LSMLIB_REAL *phi_yy,
                                    const LSMLIB_REAL *phi_yz, const LSMLIB_REAL *phi_zz,
                                    LSM_AMD_BOX(grad2_phi_gb), const LSMLIB_REAL *b,
                                    LSM_AMD_BOX(fb));
/* LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:398.
 * Computed over the reference's own loop box (its k and j loops run from the
 * fill box to the rhs ghost box, its i loop over the rhs ghost box,
 * lsm_level_set_evolution3d.f.in:620-622); every array must cover that box. */
void lsm3daddexternalandnormalveltermtolserhs_(
    LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSM_AMD_BOX(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
    const LSMLIB_REAL *phi_z_minus, LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *vel_n,
    const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
    LSM_AMD_BOX(vel_gb), LSM_AMD_BOX(rhs_fb));

/* --- normals: src/toolbox/geometry/lsm_geometry3d.h --- */
/* LSM3D_COMPUTE_UNIT_NORMAL, lsm_geometry3d.h:74 */
void lsm3dcomputeunitnormal_(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMLIB_REAL *normal_z,
                             LSM_AMD_BOX(normal_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                             const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(fb));
/* LSM3D_COMPUTE_SIGNED_UNIT_NORMAL, lsm_geometry3d.h:127 */
void lsm3dcomputesignedunitnormal_(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMLIB_REAL *normal_z,
                                   LSM_AMD_BOX(normal_gb), const LSMLIB_REAL *phi_x,
                                   const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                   LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                   LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                   const LSMLIB_REAL *dz);

/* --- field extension: src/toolbox/field_extension/lsm_field_extension3d.h --- */
/* LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS, lsm_field_extension3d.h:65 (phi read at
 * fb +- 1; S is accepted and not read, as in the reference) */
void lsm3dcomputefieldextensioneqnrhs_(LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *S,
                                       LSM_AMD_BOX(S_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                       const LSMLIB_REAL *S_x_upwind, const LSMLIB_REAL *S_y_upwind,
                                       const LSMLIB_REAL *S_z_upwind, LSM_AMD_BOX(grad_S_upwind_gb),
                                       const LSMLIB_REAL *signed_normal_x, const LSMLIB_REAL *signed_normal_y,
                                       const LSMLIB_REAL *signed_normal_z, LSM_AMD_BOX(signed_normal_gb),
                                       LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                       const LSMLIB_REAL *dz);

/* --- TVD Runge-Kutta: src/toolbox/time_integration/lsm_tvd_runge_kutta3d.h --- */
/* LSM3D_RK1_STEP, lsm_tvd_runge_kutta3d.h:59 */
void lsm3drk1step_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb), const LSMLIB_REAL *u_cur,
                   LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb),
                   LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);
/* LSM3D_TVD_RK2_STAGE1, lsm_tvd_runge_kutta3d.h:108 */
void lsm3dtvdrk2stage1_(LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                        const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs,
                        LSM_AMD_BOX(rhs_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);
/* LSM3D_TVD_RK2_STAGE2, lsm_tvd_runge_kutta3d.h:154 */
void lsm3dtvdrk2stage2_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb),
                        const LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                        const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs,
                        LSM_AMD_BOX(rhs_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);
/* LSM3D_TVD_RK3_STAGE1, lsm_tvd_runge_kutta3d.h:210 */
void lsm3dtvdrk3stage1_(LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                        const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs,
                        LSM_AMD_BOX(rhs_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);
/* LSM3D_TVD_RK3_STAGE2, lsm_tvd_runge_kutta3d.h:256 */
void lsm3dtvdrk3stage2_(LSMLIB_REAL *u_stage2, LSM_AMD_BOX(u_stage2_gb),
                        const LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                        const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs,
                        LSM_AMD_BOX(rhs_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);
/* LSM3D_TVD_RK3_STAGE3, lsm_tvd_runge_kutta3d.h:309 */
void lsm3dtvdrk3stage3_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb),
                        const LSMLIB_REAL *u_stage2, LSM_AMD_BOX(u_stage2_gb),
                        const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs,
                        LSM_AMD_BOX(rhs_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dt);

/* --- reinitialization: src/toolbox/reinitialization/lsm_reinitialization3d.h --- */
/* LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS, lsm_reinitialization3d.h:77 */
void lsm3dcomputereinitializationeqnrhs_(
    LSMLIB_REAL *reinit_rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi0, LSM_AMD_BOX(phi0_gb), const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSM_AMD_BOX(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
    const LSMLIB_REAL *phi_z_minus, LSM_AMD_BOX(grad_phi_minus_gb), LSM_AMD_BOX(fb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const int *use_phi0_for_sgn);
/* LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS, lsm_reinitialization3d.h:163 */
void lsm3dcomputeorthogonalizationeqnrhs_(
    LSMLIB_REAL *ortho_rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSM_AMD_BOX(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
    const LSMLIB_REAL *phi_z_minus, LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *psi,
    LSM_AMD_BOX(psi_gb), const LSMLIB_REAL *psi_x_plus, const LSMLIB_REAL *psi_y_plus,
    const LSMLIB_REAL *psi_z_plus, LSM_AMD_BOX(grad_psi_plus_gb),
    const LSMLIB_REAL *psi_x_minus, const LSMLIB_REAL *psi_y_minus,
    const LSMLIB_REAL *psi_z_minus, LSM_AMD_BOX(grad_psi_minus_gb), LSM_AMD_BOX(fb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_COMPUTE_DISTANCE_FOR_SUBCELL_FIX, lsm_reinitialization3d.h:247 */
void lsm3dcomputedistanceforsubcellfix_(LSMLIB_REAL *distance0, const LSMLIB_REAL *phi0,
                                        LSM_AMD_BOX(phi0_gb), LSM_AMD_BOX(fb),
                                        const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                        const LSMLIB_REAL *dz);
/* LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_SUBCELL_FIX_ORDER1, lsm_reinitialization3d.h:295 */
void lsm3dcomputereinitializationeqnrhssubcellfixorder1_(
    LSMLIB_REAL *reinit_rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi0, const LSMLIB_REAL *distance0, LSM_AMD_BOX(phi0_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,
    LSM_AMD_BOX(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus,
    LSM_AMD_BOX(grad_phi_minus_gb), LSM_AMD_BOX(fb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* --- reductions: src/toolbox/utilities/lsm_utilities3d.h, geometry/lsm_geometry3d.h --- */
/* LSM3D_MAX_NORM_DIFF, lsm_utilities3d.h:87 */
void lsm3dmaxnormdiff_(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1,
                       LSM_AMD_BOX(field1_gb), const LSMLIB_REAL *field2, LSM_AMD_BOX(field2_gb),
                       LSM_AMD_BOX(ib));
/* LSM3D_MAX_NORM_DIFF_CONTROL_VOLUME, lsm_utilities3d.h:533 */
void lsm3dmaxnormdiffcontrolvolume_(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1,
                                    LSM_AMD_BOX(field1_gb), const LSMLIB_REAL *field2,
                                    LSM_AMD_BOX(field2_gb), const LSMLIB_REAL *control_vol,
                                    LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
                                    LSM_AMD_BOX(ib));
/* LSM3D_COMPUTE_STABLE_ADVECTION_DT, lsm_utilities3d.h:221 */
void lsm3dcomputestableadvectiondt_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x,
                                    const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
                                    LSM_AMD_BOX(vel_gb), LSM_AMD_BOX(ib), const LSMLIB_REAL *dx,
                                    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                    const LSMLIB_REAL *cfl_number);
/* LSM3D_COMPUTE_STABLE_ADVECTION_DT_CONTROL_VOLUME, lsm_utilities3d.h:585 */
void lsm3dcomputestableadvectiondtcontrolvolume_(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb), const LSMLIB_REAL *control_vol,
    LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn, LSM_AMD_BOX(ib),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *cfl_number);
/* LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT, lsm_utilities3d.h:268 */
void lsm3dcomputestablenormalveldt_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                    LSM_AMD_BOX(vel_gb), const LSMLIB_REAL *phi_x_plus,
                                    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,
                                    LSM_AMD_BOX(grad_phi_plus_gb),
                                    const LSMLIB_REAL *phi_x_minus,
                                    const LSMLIB_REAL *phi_y_minus,
                                    const LSMLIB_REAL *phi_z_minus,
                                    LSM_AMD_BOX(grad_phi_minus_gb), LSM_AMD_BOX(ib),
                                    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                    const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT, lsm_utilities3d.h:331 */
void lsm3dcomputestableconstnormalveldt_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                         const LSMLIB_REAL *phi_x_plus,
                                         const LSMLIB_REAL *phi_y_plus,
                                         const LSMLIB_REAL *phi_z_plus,
                                         LSM_AMD_BOX(grad_phi_plus_gb),
                                         const LSMLIB_REAL *phi_x_minus,
                                         const LSMLIB_REAL *phi_y_minus,
                                         const LSMLIB_REAL *phi_z_minus,
                                         LSM_AMD_BOX(grad_phi_minus_gb), LSM_AMD_BOX(ib),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                         const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO, lsm_utilities3d.h:383 */
void lsm3dvolumeintegralphilessthanzero_(LSMLIB_REAL *int_F, const LSMLIB_REAL *F,
                                         LSM_AMD_BOX(F_gb), const LSMLIB_REAL *phi,
                                         LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(ib),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                         const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO, lsm_utilities3d.h:430 */
void lsm3dvolumeintegralphigreaterthanzero_(LSMLIB_REAL *int_F, const LSMLIB_REAL *F,
                                            LSM_AMD_BOX(F_gb), const LSMLIB_REAL *phi,
                                            LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(ib),
                                            const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                            const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO (interface mass), lsm_geometry3d.h:180 */
void lsm3dvolumeregionphilessthanzero_(LSMLIB_REAL *volume, const LSMLIB_REAL *phi,
                                       LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(ib),
                                       const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                       const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO, lsm_geometry3d.h:217 */
void lsm3dvolumeregionphigreaterthanzero_(LSMLIB_REAL *volume, const LSMLIB_REAL *phi,
                                          LSM_AMD_BOX(phi_gb), LSM_AMD_BOX(ib),
                                          const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                          const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);

/* LSM3D_AVE_ABS_DIFF, lsm_utilities3d.h:124 */
void lsm3daveabsdiff_(LSMLIB_REAL *ave_abs_diff, const LSMLIB_REAL *field1, LSM_AMD_BOX(field1_gb),
                      const LSMLIB_REAL *field2, LSM_AMD_BOX(field2_gb), LSM_AMD_BOX(ib));
/* LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO, lsm_utilities3d.h:160 */
void lsm3dvoxelcountgreaterthanzero_(int *count, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                     LSM_AMD_BOX(ib));
/* LSM3D_VOXEL_COUNT_LESS_THAN_ZERO, lsm_utilities3d.h:190 */
void lsm3dvoxelcountlessthanzero_(int *count, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                  LSM_AMD_BOX(ib));
/* LSM3D_SURFACE_INTEGRAL, lsm_utilities3d.h:477 */
void lsm3dsurfaceintegral_(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSM_AMD_BOX(F_gb),
                           const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x,
                           const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                           LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(ib), const LSMLIB_REAL *dx,
                           const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                           const LSMLIB_REAL *epsilon);
/* LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME, lsm_utilities3d.h:646 */
void lsm3dcomputestablenormalveldtcontrolvolume_(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSM_AMD_BOX(vel_gb), const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSM_AMD_BOX(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
    const LSMLIB_REAL *phi_z_minus, LSM_AMD_BOX(grad_phi_minus_gb),
    const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
    LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *cfl_number);
/* LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_CONTROL_VOLUME, lsm_utilities3d.h:723 */
void lsm3dcomputestableconstnormalveldtcontrolvolume_(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSM_AMD_BOX(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
    const LSMLIB_REAL *phi_z_minus, LSM_AMD_BOX(grad_phi_minus_gb),
    const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
    LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *cfl_number);
/* LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO_CONTROL_VOLUME, lsm_utilities3d.h:789 */
void lsm3dvolumeintegralphilessthanzerocontrolvolume_(
    LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSM_AMD_BOX(F_gb), const LSMLIB_REAL *phi,
    LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb),
    const int *control_vol_sgn, LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME, lsm_utilities3d.h:850 */
void lsm3dvolumeintegralphigreaterthanzerocontrolvolume_(
    LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSM_AMD_BOX(F_gb), const LSMLIB_REAL *phi,
    LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb),
    const int *control_vol_sgn, LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_SURFACE_INTEGRAL_CONTROL_VOLUME, lsm_utilities3d.h:911 */
void lsm3dsurfaceintegralcontrolvolume_(
    LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSM_AMD_BOX(F_gb), const LSMLIB_REAL *phi,
    LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
    const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *control_vol,
    LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn, LSM_AMD_BOX(ib),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *epsilon);
/* LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO_CONTROL_VOLUME, lsm_utilities3d.h:969.
 * Counts cells with phi > 0 and control_vol >= 0 (sign +1) or <= 0 (sign
 * -1): unlike the other control-volume routines, cells with control_vol
 * equal to zero are included (lsm_utilities3d.f.in:1895-1896). */
void lsm3dvoxelcountgreaterthanzerocontrolvolume_(int *count, const LSMLIB_REAL *phi,
                                                  LSM_AMD_BOX(phi_gb),
                                                  const LSMLIB_REAL *control_vol,
                                                  LSM_AMD_BOX(control_vol_gb),
                                                  const int *control_vol_sgn, LSM_AMD_BOX(ib));
/* LSM3D_VOXEL_COUNT_LESS_THAN_ZERO_CONTROL_VOLUME, lsm_utilities3d.h:1006 (zero
 * control_vol included as above) */
void lsm3dvoxelcountlessthanzerocontrolvolume_(int *count, const LSMLIB_REAL *phi,
                                               LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *control_vol,
                                               LSM_AMD_BOX(control_vol_gb),
                                               const int *control_vol_sgn, LSM_AMD_BOX(ib));
/* LSM3D_SURFACE_AREA_ZERO_LEVEL_SET, lsm_geometry3d.h:256 */
void lsm3dsurfaceareazerolevelset_(LSMLIB_REAL *area, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                   const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                   const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
                                   LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                   const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA, lsm_geometry3d.h:301 */
void lsm3dsurfaceareazerolevelsetdelta_(LSMLIB_REAL *area, const LSMLIB_REAL *delta_phi,
                                        LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x,
                                        const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                        LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(ib),
                                        const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                        const LSMLIB_REAL *dz);
/* LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO_CONTROL_VOLUME, lsm_geometry3d.h:351 */
void lsm3dvolumeregionphilessthanzerocontrolvolume_(
    LSMLIB_REAL *volume, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
    LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *epsilon);
/* LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME, lsm_geometry3d.h:402 */
void lsm3dvolumeregionphigreaterthanzerocontrolvolume_(
    LSMLIB_REAL *volume, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
    LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *epsilon);
/* LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_CONTROL_VOLUME, lsm_geometry3d.h:455 */
void lsm3dsurfaceareazerolevelsetcontrolvolume_(
    LSMLIB_REAL *area, const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x,
    const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
    const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb), const int *control_vol_sgn,
    LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *epsilon);
/* LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA_CONTROL_VOLUME, lsm_geometry3d.h:514 */
void lsm3dsurfaceareazerolevelsetdeltacontrolvolume_(
    LSMLIB_REAL *area, const LSMLIB_REAL *delta_phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
    LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *control_vol, LSM_AMD_BOX(control_vol_gb),
    const int *control_vol_sgn, LSM_AMD_BOX(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const LSMLIB_REAL *dz);

/* --- delta function: src/toolbox/utilities/lsm_calculus_toolbox3d.h --- */
/* LSM3D_DELTA_FUNCTION_ORDER1, lsm_calculus_toolbox3d.h:47.  The reference
 * header declares delta const although the routine writes it; phi needs one
 * ghost cell around the fill box (the routine reads its six neighbours).  The
 * z terms use norm_phi_x and dx as the reference computes them; norm_phi_z and
 * dz are accepted and unused. */
/* qualifiers as the reference header declares them */
void lsm3ddeltafunctionorder1_(const LSMLIB_REAL *phi, const LSMLIB_REAL *delta, LSM_AMD_BOX(gb),
                               const LSMLIB_REAL *norm_phi_x, const LSMLIB_REAL *norm_phi_y,
                               const LSMLIB_REAL *norm_phi_z, LSM_AMD_BOX(grad_phi_gb),
                               LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                               const LSMLIB_REAL *dz);

/* --- curvature: src/toolbox/geometry/lsm_curvature3d.h --- */
/* Mean curvature (numerator over |grad phi|^3) and Gaussian curvature
 * (over |grad phi|^4) from phi and its precomputed first derivatives; 0 where
 * that power of grad_phi_mag is below zero_tol.  Order 2 reads phi one cell
 * around the fill box (diagonals included); order 4 reads phi two cells out
 * along each axis and the gradients two cells out (phi_x along y, phi_y along
 * z, phi_z along x). */
/* LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2, lsm_curvature3d.h:68 */
void lsm3dcomputemeancurvatureorder2_(LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi,
                                      LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x,
                                      const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                      const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb),
                                      LSM_AMD_BOX(kappa_fb), const LSMLIB_REAL *dx,
                                      const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2, lsm_curvature3d.h:128 */
void lsm3dcomputegaussiancurvatureorder2_(LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb),
                                          const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                          const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                          const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                          LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(kappa_fb),
                                          const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                          const LSMLIB_REAL *dz);
/* LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4, lsm_curvature3d.h:183 */
void lsm3dcomputemeancurvatureorder4_(LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi,
                                      LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *phi_x,
                                      const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                      const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb),
                                      LSM_AMD_BOX(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                      const LSMLIB_REAL *dz);
/* LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4, lsm_curvature3d.h:244 */
void lsm3dcomputegaussiancurvatureorder4_(LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb),
                                          const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                                          const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                          const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                          LSM_AMD_BOX(grad_phi_gb), LSM_AMD_BOX(kappa_fb),
                                          const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                          const LSMLIB_REAL *dz);
/* LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS, lsm_level_set_evolution3d.h:334.
 * lse_rhs += b kappa grad_phi_mag over the cells the reference loops over:
 * i over the whole rhs ghost box, j and k from the fill box's lower end to
 * the rhs ghost box's upper end (lsm_level_set_evolution3d.f.in:514-516);
 * kappa and grad_phi_mag must cover those cells.  The reference header
 * declares lse_rhs const and kappa non-const; the routine writes lse_rhs
 * and only reads kappa. */
/* qualifiers as the reference header declares them */
void lsm3daddconstprecomputedcurvtermtolserhs_(const LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(rhs_gb),
                                               LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb),
                                               const LSMLIB_REAL *grad_phi_mag,
                                               LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *b,
                                               LSM_AMD_BOX(rhs_fb));

/* --- box copy: src/parallel/fortran/lsm_samrai_f77_utilities.h --- */
/* LSM3D_SAMRAI_UTILITIES_COPY_DATA, lsm_samrai_f77_utilities.h:52: dst = src on
 * the fill box (the reference header declares src non-const; it is only read) */
void lsm3dsamraiutilitiescopydata_(LSMLIB_REAL *dst_data, LSM_AMD_BOX(dst_gb),
                                   const LSMLIB_REAL *src_data, LSM_AMD_BOX(src_gb),
                                   LSM_AMD_BOX(fb));

/* --- ghost fill: src/toolbox/boundary_conditions/lsm_boundary_conditions3d.h --- */
/* bdry_location_idx: 0 x_lo, 1 x_hi, 2 y_lo, 3 y_hi, 4 z_lo, 5 z_hi */
/* LSM3D_LINEAR_EXTRAPOLATION, lsm_boundary_conditions3d.h:71 */
void lsm3dlinearextrapolation_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                               const int *bdry_location_idx);
/* LSM3D_SIGNED_LINEAR_EXTRAPOLATION, lsm_boundary_conditions3d.h:111 */
void lsm3dsignedlinearextrapolation_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                                     const int *bdry_location_idx);
/* LSM3D_COPY_EXTRAPOLATION, lsm_boundary_conditions3d.h:135 */
void lsm3dcopyextrapolation_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                             const int *bdry_location_idx);
/* LSM3D_HOMOGENEOUS_NEUMANN_ENO1, lsm_boundary_conditions3d.h:161 */
void lsm3dhomogeneousneumanneno1_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                                  const int *bdry_location_idx);
/* LSM3D_HOMOGENEOUS_NEUMANN_ENO2, lsm_boundary_conditions3d.h:186 */
void lsm3dhomogeneousneumanneno2_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                                  const int *bdry_location_idx);
/* LSM3D_HOMOGENEOUS_NEUMANN_ENO3, lsm_boundary_conditions3d.h:211 */
void lsm3dhomogeneousneumanneno3_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                                  const int *bdry_location_idx);
/* LSM3D_HOMOGENEOUS_NEUMANN_WENO5, lsm_boundary_conditions3d.h:236 */
void lsm3dhomogeneousneumannweno5_(LSMLIB_REAL *phi, LSM_AMD_BOX(gb), LSM_AMD_BOX(fb),
                                   const int *bdry_location_idx);

/* --- narrow band (Peng et al.'s local level set method) ---------------------
 * Index lists index_x/y/z(nlo:nhi): entry l at element l - nlo of the array
 * passed.  The lists, index_outer and narrow_band may be device or host
 * memory; n_lo / n_hi, change_sign and the other int outputs are host
 * memory.  Marks compare as signed bytes (the Fortran's integer*1).  List
 * entries outside the arrays they index are skipped and reported as an error
 * (undefined behaviour in the reference).  HJ_ENO1/2_LOCAL write D1/D2 as
 * the reference does.  COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL uses the box
 * routine's |vel_n| and smoothed |grad phi| (the reference reads two
 * uninitialised variables there, lsm_utilities3d_local.f.in:728).  Inputs
 * the reference headers declare non-const (and the outputs of COPY_DATA_LOCAL
 * and GRADIENT_MAGNITUDE_LOCAL, which they declare const) are declared here
 * by how the routines use them; the symbols and argument order are theirs. */
/* --- narrow band: src/toolbox/localization/lsm_localization3d.h --- */
/* LSM3D_DETERMINE_NARROW_BAND, lsm_localization3d.h:89 */
void lsm3ddeterminenarrowband_(const LSMLIB_REAL *phi, LSM_AMD_BOX(gb), unsigned char *narrow_band,
                               LSM_AMD_BOX(nb_gb), int *index_x, int *index_y, int *index_z,
                               const int *nlo_index, const int *nhi_index, int *n_lo, int *n_hi,
                               int *index_outer, const int *nlo_index_outer,
                               const int *nhi_index_outer, int *nlo_index_outer_plus,
                               int *nhi_index_outer_plus, int *nlo_index_outer_minus,
                               int *nhi_index_outer_minus, const LSMLIB_REAL *width,
                               const LSMLIB_REAL *width_inner, const int *level);
/* LSM3D_MARK_NARROW_BAND_BOUNDARY_LAYER, lsm_localization3d.h:138 */
void lsm3dmarknarrowbandboundarylayer_(unsigned char *narrow_band, LSM_AMD_BOX(gb),
                                       LSM_AMD_BOX(fb), const unsigned char *mark_boundary_layer);
/* LSM3D_DETERMINE_NARROW_BAND_FROM_MASK, lsm_localization3d.h:187 */
void lsm3ddeterminenarrowbandfrommask_(const LSMLIB_REAL *mask, LSM_AMD_BOX(mask_gb),
                                       unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                       int *index_x, int *index_y, int *index_z, int *nlo_index,
                                       int *nhi_index, int *n_lo, int *n_hi, const int *level,
                                       const int *use_mask_sign);
/* LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL, lsm_localization3d.h:230 */
void lsm3dmultiplycutofflserhslocal_(const LSMLIB_REAL *phi, LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(gb),
                                     const int *index_x, const int *index_y, const int *index_z,
                                     const int *nlo_index, const int *nhi_index,
                                     const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                     const unsigned char *mark_fb, const LSMLIB_REAL *beta,
                                     const LSMLIB_REAL *gamma);
/* LSM3D_CHECK_OUTER_NARROW_BAND_LAYER, lsm_localization3d.h:270 */
/* qualifiers as the reference header declares them */
void lsm3dcheckouternarrowbandlayer_(int *change_sign, const LSMLIB_REAL *phi, LSM_AMD_BOX(gb),
                                     int *index_x, int *index_y, int *index_z,
                                     int *nlo_index, int *nhi_index, const int *index_outer,
                                     const int *nlo_index_outer, const int *nhi_index_outer,
                                     const int *nlo_index_outer_plus,
                                     const int *nhi_index_outer_plus,
                                     const int *nlo_index_outer_minus,
                                     const int *nhi_index_outer_minus);
/* LSM3D_IMPOSE_MASK_LOCAL, lsm_localization3d.h:307 */
/* qualifiers as the reference header declares them */
void lsm3dimposemasklocal_(LSMLIB_REAL *dest, LSM_AMD_BOX(dest_gb), LSMLIB_REAL *src,
                           LSM_AMD_BOX(src_gb), LSMLIB_REAL *mask, LSM_AMD_BOX(mask_gb),
                           const int *index_x, const int *index_y, const int *index_z,
                           const int *nlo_index, const int *nhi_index);
/* LSM3D_COPY_DATA_LOCAL, lsm_localization3d.h:347 */
/* qualifiers as the reference header declares them */
void lsm3dcopydatalocal_(const LSMLIB_REAL *dest, LSM_AMD_BOX(dest_gb), const LSMLIB_REAL *src,
                         LSM_AMD_BOX(src_gb), const int *index_x, const int *index_y,
                         const int *index_z, const int *nlo_index, const int *nhi_index);

/* --- narrow-band kernels: src/toolbox/spatial_derivatives/lsm_spatial_derivatives3d_local.h --- */
/* LSM3D_HJ_ENO1_LOCAL, lsm_spatial_derivatives3d_local.h:78 */
void lsm3dhjeno1local_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                       LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                       LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                       LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                       LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), const LSMLIB_REAL *dx,
                       const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x,
                       const int *index_y, const int *index_z, const int *nlo_index0,
                       const int *nhi_index0, const int *nlo_index1, const int *nhi_index1,
                       const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                       const unsigned char *mark_fb, const unsigned char *mark_D1);
/* LSM3D_HJ_ENO2_LOCAL, lsm_spatial_derivatives3d_local.h:159 */
void lsm3dhjeno2local_(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                       LSM_AMD_BOX(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus,
                       LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,
                       LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
                       LSMLIB_REAL *D1, LSM_AMD_BOX(D1_gb), LSMLIB_REAL *D2, LSM_AMD_BOX(D2_gb),
                       const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                       const int *index_x, const int *index_y, const int *index_z,
                       const int *nlo_index0, const int *nhi_index0, const int *nlo_index1,
                       const int *nhi_index1, const int *nlo_index2, const int *nhi_index2,
                       const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                       const unsigned char *mark_fb, const unsigned char *mark_D1,
                       const unsigned char *mark_D2);
/* LSM3D_CENTRAL_GRAD_ORDER2_LOCAL, lsm_spatial_derivatives3d_local.h:240 */
void lsm3dcentralgradorder2local_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                                  LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi,
                                  LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *dx,
                                  const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x,
                                  const int *index_y, const int *index_z, const int *nlo_index,
                                  const int *nhi_index, const unsigned char *narrow_band,
                                  LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_CENTRAL_GRAD_ORDER4_LOCAL, lsm_spatial_derivatives3d_local.h:293 */
void lsm3dcentralgradorder4local_(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                                  LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *phi,
                                  LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *dx,
                                  const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x,
                                  const int *index_y, const int *index_z, const int *nlo_index,
                                  const int *nhi_index, const unsigned char *narrow_band,
                                  LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_LAPLACIAN_ORDER2_LOCAL, lsm_spatial_derivatives3d_local.h:345 */
void lsm3dlaplacianorder2local_(LSMLIB_REAL *laplacian_phi, LSM_AMD_BOX(laplacian_phi_gb),
                                const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *dx,
                                const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x,
                                const int *index_y, const int *index_z, const int *nlo_index,
                                const int *nhi_index, const unsigned char *narrow_band,
                                LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL, lsm_spatial_derivatives3d_local.h:397 */
void lsm3dcomputeavegradphilocal_(LSMLIB_REAL *grad_phi_ave, const LSMLIB_REAL *phi,
                                  LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *dx,
                                  const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x,
                                  const int *index_y, const int *index_z, const int *nlo_index,
                                  const int *nhi_index, const unsigned char *narrow_band,
                                  LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_GRADIENT_MAGNITUDE_LOCAL, lsm_spatial_derivatives3d_local.h:438 (the
 * reference declares the output grad_phi_mag const; kept, and written) */
void lsm3dgradientmagnitudelocal_(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                  const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                  LSM_AMD_BOX(grad_phi_gb), const int *index_x, const int *index_y,
                                  const int *index_z, const int *nlo_index, const int *nhi_index,
                                  const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                  const unsigned char *mark_fb);

/* --- narrow-band kernels: src/toolbox/time_integration/lsm_tvd_runge_kutta3d_local.h --- */
/* LSM3D_RK1_STEP_LOCAL, lsm_tvd_runge_kutta3d_local.h:35 */
void lsm3drk1steplocal_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb), const LSMLIB_REAL *u_cur,
                        LSM_AMD_BOX(u_cur_gb), const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb),
                        const LSMLIB_REAL *dt, const int *index_x, const int *index_y,
                        const int *index_z, const int *nlo_index, const int *nhi_index,
                        const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                        const unsigned char *mark_fb);
/* LSM3D_TVD_RK2_STAGE1_LOCAL, lsm_tvd_runge_kutta3d_local.h:73 */
void lsm3dtvdrk2stage1local_(LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                             const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb),
                             const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *dt,
                             const int *index_x, const int *index_y, const int *index_z,
                             const int *nlo_index, const int *nhi_index,
                             const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                             const unsigned char *mark_fb);
/* LSM3D_TVD_RK2_STAGE2_LOCAL, lsm_tvd_runge_kutta3d_local.h:111 */
void lsm3dtvdrk2stage2local_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb),
                             const LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                             const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb),
                             const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *dt,
                             const int *index_x, const int *index_y, const int *index_z,
                             const int *nlo_index, const int *nhi_index,
                             const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                             const unsigned char *mark_fb);
/* LSM3D_TVD_RK3_STAGE1_LOCAL, lsm_tvd_runge_kutta3d_local.h:155 */
void lsm3dtvdrk3stage1local_(LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                             const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb),
                             const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *dt,
                             const int *index_x, const int *index_y, const int *index_z,
                             const int *nlo_index, const int *nhi_index,
                             const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                             const unsigned char *mark_fb);
/* LSM3D_TVD_RK3_STAGE2_LOCAL, lsm_tvd_runge_kutta3d_local.h:192 */
void lsm3dtvdrk3stage2local_(LSMLIB_REAL *u_stage2, LSM_AMD_BOX(u_stage2_gb),
                             const LSMLIB_REAL *u_stage1, LSM_AMD_BOX(u_stage1_gb),
                             const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb),
                             const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *dt,
                             const int *index_x, const int *index_y, const int *index_z,
                             const int *nlo_index, const int *nhi_index,
                             const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                             const unsigned char *mark_fb);
/* LSM3D_TVD_RK3_STAGE3_LOCAL, lsm_tvd_runge_kutta3d_local.h:237 */
void lsm3dtvdrk3stage3local_(LSMLIB_REAL *u_next, LSM_AMD_BOX(u_next_gb),
                             const LSMLIB_REAL *u_stage2, LSM_AMD_BOX(u_stage2_gb),
                             const LSMLIB_REAL *u_cur, LSM_AMD_BOX(u_cur_gb),
                             const LSMLIB_REAL *rhs, LSM_AMD_BOX(rhs_gb), const LSMLIB_REAL *dt,
                             const int *index_x, const int *index_y, const int *index_z,
                             const int *nlo_index, const int *nhi_index,
                             const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                             const unsigned char *mark_fb);

/* --- narrow-band kernels: src/toolbox/level_set_evolution/lsm_level_set_evolution3d_local.h --- */
/* LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL, lsm_level_set_evolution3d_local.h:67 */
void lsm3dzerooutlevelseteqnrhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                      const int *index_x, const int *index_y, const int *index_z,
                                      const int *nlo_index, const int *nhi_index);
/* LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:102 */
void lsm3daddadvectiontermtolserhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                         const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
                                         const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                         const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb),
                                         const int *index_x, const int *index_y,
                                         const int *index_z, const int *nlo_index,
                                         const int *nhi_index, const unsigned char *narrow_band,
                                         LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:166 */
void lsm3daddnormalveltermtolserhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                         const LSMLIB_REAL *phi_x_plus,
                                         const LSMLIB_REAL *phi_y_plus,
                                         const LSMLIB_REAL *phi_z_plus,
                                         LSM_AMD_BOX(grad_phi_plus_gb),
                                         const LSMLIB_REAL *phi_x_minus,
                                         const LSMLIB_REAL *phi_y_minus,
                                         const LSMLIB_REAL *phi_z_minus,
                                         LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *vel_n,
                                         LSM_AMD_BOX(vel_gb), const int *index_x,
                                         const int *index_y, const int *index_z,
                                         const int *nlo_index, const int *nhi_index,
                                         const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                         const unsigned char *mark_fb);
/* LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:237 */
void lsm3daddconstnormalveltermtolserhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                              const LSMLIB_REAL *phi_x_plus,
                                              const LSMLIB_REAL *phi_y_plus,
                                              const LSMLIB_REAL *phi_z_plus,
                                              LSM_AMD_BOX(grad_phi_plus_gb),
                                              const LSMLIB_REAL *phi_x_minus,
                                              const LSMLIB_REAL *phi_y_minus,
                                              const LSMLIB_REAL *phi_z_minus,
                                              LSM_AMD_BOX(grad_phi_minus_gb),
                                              const LSMLIB_REAL *vel_n, const int *index_x,
                                              const int *index_y, const int *index_z,
                                              const int *nlo_index, const int *nhi_index,
                                              const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                              const unsigned char *mark_fb);
/* LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:300 */
void lsm3daddconstcurvtermtolserhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(lse_rhs_gb),
                                         const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, LSM_AMD_BOX(grad_phi_gb),
                                         const LSMLIB_REAL *phi_xx, const LSMLIB_REAL *phi_xy,
                                         const LSMLIB_REAL *phi_xz, const LSMLIB_REAL *phi_yy,
                                         const LSMLIB_REAL *phi_yz, const LSMLIB_REAL *phi_zz,
                                         LSM_AMD_BOX(grad2_phi_gb), const LSMLIB_REAL *b,
                                         const int *index_x, const int *index_y,
                                         const int *index_z, const int *nlo_index,
                                         const int *nhi_index, const unsigned char *narrow_band,
                                         LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:374
 * (the reference declares lse_rhs const and kappa non-const; kept, and
 * lse_rhs is updated) */
void lsm3daddconstprecomputedcurvtermtolserhslocal_(const LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(rhs_gb),
                                                    LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb),
                                                    const LSMLIB_REAL *grad_mag_phi,
                                                    LSM_AMD_BOX(phi_gb), const LSMLIB_REAL *b,
                                                    const int *index_x, const int *index_y,
                                                    const int *index_z, const int *nlo_index,
                                                    const int *nhi_index,
                                                    const unsigned char *narrow_band,
                                                    LSM_AMD_BOX(nb_gb),
                                                    const unsigned char *mark_fb);
/* LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL, lsm_level_set_evolution3d_local.h:449 */
void lsm3daddexternalandnormalveltermtolserhslocal_(LSMLIB_REAL *lse_rhs, LSM_AMD_BOX(rhs_gb),
                                                    const LSMLIB_REAL *phi_x_plus,
                                                    const LSMLIB_REAL *phi_y_plus,
                                                    const LSMLIB_REAL *phi_z_plus,
                                                    LSM_AMD_BOX(grad_phi_plus_gb),
                                                    const LSMLIB_REAL *phi_x_minus,
                                                    const LSMLIB_REAL *phi_y_minus,
                                                    const LSMLIB_REAL *phi_z_minus,
                                                    LSM_AMD_BOX(grad_phi_minus_gb),
                                                    const LSMLIB_REAL *vel_n,
                                                    const LSMLIB_REAL *vel_x,
                                                    const LSMLIB_REAL *vel_y,
                                                    const LSMLIB_REAL *vel_z, LSM_AMD_BOX(vel_gb),
                                                    const int *index_x, const int *index_y,
                                                    const int *index_z, const int *nlo_index,
                                                    const int *nhi_index,
                                                    const unsigned char *narrow_band,
                                                    LSM_AMD_BOX(nb_gb),
                                                    const unsigned char *mark_fb);

/* --- narrow-band kernels: src/toolbox/reinitialization/lsm_reinitialization3d_local.h --- */
/* LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL, lsm_reinitialization3d_local.h:41 */
void lsm3dcomputereinitializationeqnrhslocal_(LSMLIB_REAL *reinit_rhs, LSM_AMD_BOX(rhs_gb),
                                              const LSMLIB_REAL * phi, LSM_AMD_BOX(phi_gb),
                                              const LSMLIB_REAL * phi0, LSM_AMD_BOX(phi0_gb),
                                              const LSMLIB_REAL *phi_x_plus,
                                              const LSMLIB_REAL *phi_y_plus,
                                              const LSMLIB_REAL *phi_z_plus,
                                              LSM_AMD_BOX(grad_phi_plus_gb),
                                              const LSMLIB_REAL *phi_x_minus,
                                              const LSMLIB_REAL *phi_y_minus,
                                              const LSMLIB_REAL *phi_z_minus,
                                              LSM_AMD_BOX(grad_phi_minus_gb),
                                              const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                              const LSMLIB_REAL *dz, const int *use_phi0_for_sgn,
                                              const int *index_x, const int *index_y,
                                              const int *index_z, const int *nlo_index,
                                              const int *nhi_index,
                                              const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                              const unsigned char *mark_fb);
/* LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL, lsm_reinitialization3d_local.h:99 */
void lsm3dcomputeorthogonalizationeqnrhslocal_(LSMLIB_REAL *ortho_rhs, LSM_AMD_BOX(rhs_gb),
                                               const LSMLIB_REAL *phi_x_plus,
                                               const LSMLIB_REAL *phi_y_plus,
                                               const LSMLIB_REAL *phi_z_plus,
                                               LSM_AMD_BOX(grad_phi_plus_gb),
                                               const LSMLIB_REAL *phi_x_minus,
                                               const LSMLIB_REAL *phi_y_minus,
                                               const LSMLIB_REAL *phi_z_minus,
                                               LSM_AMD_BOX(grad_phi_minus_gb),
                                               const LSMLIB_REAL * psi, LSM_AMD_BOX(psi_gb),
                                               const LSMLIB_REAL *psi_x_plus,
                                               const LSMLIB_REAL *psi_y_plus,
                                               const LSMLIB_REAL *psi_z_plus,
                                               LSM_AMD_BOX(grad_psi_plus_gb),
                                               const LSMLIB_REAL *psi_x_minus,
                                               const LSMLIB_REAL *psi_y_minus,
                                               const LSMLIB_REAL *psi_z_minus,
                                               LSM_AMD_BOX(grad_psi_minus_gb),
                                               const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                               const LSMLIB_REAL *dz, const int *index_x,
                                               const int *index_y, const int *index_z,
                                               const int *nlo_index, const int *nhi_index,
                                               const unsigned char *narrow_band,
                                               LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);

/* --- narrow-band kernels: src/toolbox/utilities/lsm_utilities3d_local.h --- */
/* LSM3D_MAX_NORM_DIFF_LOCAL, lsm_utilities3d_local.h:65 */
void lsm3dmaxnormdifflocal_(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1,
                            LSM_AMD_BOX(field1_gb), const LSMLIB_REAL *field2,
                            LSM_AMD_BOX(field2_gb), const int *index_x, const int *index_y,
                            const int *index_z, const int *nlo_index, const int *nhi_index,
                            const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                            const unsigned char *mark_fb);
/* LSM3D_COMPUTE_STABLE_ADVECTION_DT_LOCAL, lsm_utilities3d_local.h:112 */
void lsm3dcomputestableadvectiondtlocal_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x,
                                         const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
                                         LSM_AMD_BOX(vel_gb), const LSMLIB_REAL *dx,
                                         const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                         const LSMLIB_REAL *cfl_number, const int *index_x,
                                         const int *index_y, const int *index_z,
                                         const int *nlo_index, const int *nhi_index,
                                         const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                         const unsigned char *mark_fb);
/* LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL, lsm_utilities3d_local.h:169 */
void lsm3dcomputestablenormalveldtlocal_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                         LSM_AMD_BOX(vel_gb), const LSMLIB_REAL *phi_x_plus,
                                         const LSMLIB_REAL *phi_y_plus,
                                         const LSMLIB_REAL *phi_z_plus,
                                         LSM_AMD_BOX(grad_phi_plus_gb),
                                         const LSMLIB_REAL *phi_x_minus,
                                         const LSMLIB_REAL *phi_y_minus,
                                         const LSMLIB_REAL *phi_z_minus,
                                         LSM_AMD_BOX(grad_phi_minus_gb), const LSMLIB_REAL *dx,
                                         const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                         const LSMLIB_REAL *cfl_number, const int *index_x,
                                         const int *index_y, const int *index_z,
                                         const int *nlo_index, const int *nhi_index,
                                         const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                         const unsigned char *mark_fb);
/* LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME_LOCAL, lsm_utilities3d_local.h:241 */
void lsm3dcomputestablenormalveldtcontrolvolumelocal_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                                      LSM_AMD_BOX(vel_gb),
                                                      const LSMLIB_REAL *phi_x_plus,
                                                      const LSMLIB_REAL *phi_y_plus,
                                                      const LSMLIB_REAL *phi_z_plus,
                                                      LSM_AMD_BOX(grad_phi_plus_gb),
                                                      const LSMLIB_REAL *phi_x_minus,
                                                      const LSMLIB_REAL *phi_y_minus,
                                                      const LSMLIB_REAL *phi_z_minus,
                                                      LSM_AMD_BOX(grad_phi_minus_gb),
                                                      const LSMLIB_REAL *control_vol,
                                                      LSM_AMD_BOX(control_vol_gb),
                                                      const int *control_vol_sgn,
                                                      const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                                      const LSMLIB_REAL *dz,
                                                      const LSMLIB_REAL *cfl_number,
                                                      const int *index_x, const int *index_y,
                                                      const int *index_z, const int *nlo_index,
                                                      const int *nhi_index,
                                                      const unsigned char *narrow_band,
                                                      LSM_AMD_BOX(nb_gb),
                                                      const unsigned char *mark_fb);
/* LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL, lsm_utilities3d_local.h:322 */
void lsm3dcomputestableconstnormalveldtlocal_(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                              const LSMLIB_REAL *phi_x_plus,
                                              const LSMLIB_REAL *phi_y_plus,
                                              const LSMLIB_REAL *phi_z_plus,
                                              LSM_AMD_BOX(grad_phi_plus_gb),
                                              const LSMLIB_REAL *phi_x_minus,
                                              const LSMLIB_REAL *phi_y_minus,
                                              const LSMLIB_REAL *phi_z_minus,
                                              LSM_AMD_BOX(grad_phi_minus_gb),
                                              const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                              const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number,
                                              const int *index_x, const int *index_y,
                                              const int *index_z, const int *nlo_index,
                                              const int *nhi_index,
                                              const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb),
                                              const unsigned char *mark_fb);

/* --- narrow-band kernels: src/toolbox/geometry/lsm_curvature3d_local.h --- */
/* As the box routines above, at the listed points marked <= mark_fb. */
/* LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL, lsm_curvature3d_local.h:72 */
void lsm3dcomputemeancurvatureorder2local_(
    LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
    const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x, const int *index_y,
    const int *index_z, const int *nlo_index, const int *nhi_index,
    const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2_LOCAL, lsm_curvature3d_local.h:141 */
void lsm3dcomputegaussiancurvatureorder2local_(
    LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
    const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x, const int *index_y,
    const int *index_z, const int *nlo_index, const int *nhi_index,
    const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL, lsm_curvature3d_local.h:210.
 * The reference's loop body computes the fourth-order GAUSSIAN curvature
 * (numerator and |grad phi|^4, lsm_curvature3d_local.f.in:403-417); this
 * drop-in returns the same. */
void lsm3dcomputemeancurvatureorder4local_(
    LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
    const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x, const int *index_y,
    const int *index_z, const int *nlo_index, const int *nhi_index,
    const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);
/* LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4_LOCAL, lsm_curvature3d_local.h:279 */
void lsm3dcomputegaussiancurvatureorder4local_(
    LSMLIB_REAL *kappa, LSM_AMD_BOX(kappa_gb), const LSMLIB_REAL *phi, LSM_AMD_BOX(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
    const LSMLIB_REAL *grad_phi_mag, LSM_AMD_BOX(grad_phi_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const int *index_x, const int *index_y,
    const int *index_z, const int *nlo_index, const int *nhi_index,
    const unsigned char *narrow_band, LSM_AMD_BOX(nb_gb), const unsigned char *mark_fb);

/* ======================================================================
 * Part 2: runtime control (no counterpart in the reference: the Fortran
 * library has no stream, no error channel and no device)
 * ====================================================================== */
/* Stream (a hipStream_t) that this host thread's calls queue work on; NULL =
 * the default stream. */
void lsm_amd_set_stream(void *stream);
void *lsm_amd_get_stream(void);
/* 1 (default): Part 1 calls complete before returning, as the Fortran does.
 * 0: they only order work on the stream (host-memory and reduction calls
 * still wait, since they must hand back host data). */
void lsm_amd_set_sync(int sync);
/* 0 (default): invalid arguments or HIP errors print a message and abort.
 * 1: the message is recorded for lsm_amd_last_error() and the call returns. */
void lsm_amd_set_error_mode(int mode);
int lsm_amd_error_pending(void);
const char *lsm_amd_last_error(void);
void lsm_amd_clear_error(void);
/* sizeof(LSMLIB_REAL) of this library (8: double, 4: float) */
int lsm_amd_real_size(void);
/* Streaming-bandwidth probe (a measurement utility; no counterpart in the
 * reference): `reps` launches of a memory-only kernel over device arrays on
 * the library stream, timed with HIP events; returns milliseconds per launch
 * (-1 on error).  16-byte non-temporal accesses, `workgroups` x 256 threads.
 *   kind 0: copy, dst[0:bytes) = src[0:bytes)
 *   kind 1: 5 reads + 1 write, dst = sum of the five consecutive `bytes`
 *           slices of src (src holds 5 * bytes)
 *   kind 2: reads of src[0:bytes); one float per thread is written to dst
 *           (dst holds workgroups * 256 floats)
 *   kind 3: as kind 1 with 4-byte accesses (one float per lane)
 * bytes must be a positive multiple of 16. */
double lsm_amd_bandwidth_probe(int kind, const void *src, void *dst, long long bytes, int workgroups,
                               int reps);
/* The stage kernels' tile march with no arithmetic (a measurement utility;
 * no counterpart in the reference): 32 x 8 column tiles marching over z-chunks
 * of kz planes in the stage launch's tile order, reading 5 fields and writing
 * 1 at every interior cell of an n^3 box, non-temporal 4-byte accesses, no
 * halo.  src holds 5 consecutive fields and dst 1, each px * py * (n + 6)
 * floats (3 ghost layers per side; interior cell (0,0,0) at row offset 32 of
 * row 3 of plane 3).  n is a multiple of 64 and of kz, px >= n + 64,
 * py >= n + 6.  Returns milliseconds per launch (-1 on error). */
double lsm_amd_march_probe(const void *src, void *dst, int n, int px, int py, int kz, int reps);
/* FP32 vector-ALU issue-rate probe (a measurement utility; no counterpart in
 * the reference): `workgroups` x 256 lanes each run trips x 32 instructions in
 * eight independent chains, no memory traffic in the loop: op 0 v_fma_f32 with
 * scalar-register operands, 1 the transcendental v_rcp_f32, 2 v_mul_f32 and
 * 3 v_add_f32 with a scalar-register operand, 4 v_fma_f32 with per-lane
 * operands, 5 the accumulating v_fmac_f32, 6 v_max3_f32, 7 v_fmamk_f32 with
 * literal constants; dst
 * receives one float per lane (workgroups * 256 floats).  Returns
 * milliseconds per launch (-1 on error); lane-instructions per second =
 * workgroups * 256 * trips * 32 / time. */
double lsm_amd_valu_probe(void *dst, int op, int workgroups, int trips, int reps);

/* ======================================================================
 * Part 3: fused device-resident stages (this library's own API; no
 * counterpart in the reference).  All arrays must be device memory; work is
 * queued on the library stream; calls return at once under
 * lsm_amd_set_sync(0).  Boxes are int[6] = {ilo, ihi, jlo, jhi, klo, khi}.
 * ====================================================================== */
/* Runge-Kutta forms: phi_out = ...
 *   1  phi_cur + dt*rhs                       (RK1, TVD-RK2/3 stage 1)
 *   22 0.5*(phi_cur + phi_stage + dt*rhs)     (TVD-RK2 stage 2)
 *   32 0.75*phi_cur + 0.25*(phi_stage + dt*rhs)   (TVD-RK3 stage 2)
 *   33 phi_cur/3 + 2/3*(phi_stage + dt*rhs)   (TVD-RK3 stage 3)
 * phi_cur may be NULL for form 1 (then phi_cur = phi_stage).  phi_out may
 * alias phi_cur (it is read only at the cell being written) but not
 * phi_stage.
 * bc[6] (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi), or NULL: how the stage obtains
 * the ghost values of phi_stage it reads.  0 = read the stored ghost cells;
 * 1 copy (homogeneous Neumann), 2 linear, 3 signed linear extrapolation, 5
 * periodic = synthesised from the fill box in the kernel, with exactly the
 * values LSM3D_*_EXTRAPOLATION (or a periodic fill) would have stored, so no
 * ghost fill is needed between stages.
 * periodic_mask bit d (x=1, y=2, z=4): the stage also writes phi_out's
 * periodic ghost images along axis d (3-cell ghost layer), for callers that
 * read phi_out's ghost cells afterwards.
 * periodic_mask bit 8 (LSM_AMD_WRITE_FACE_GHOSTS): bc[6] then describes
 * phi_out's faces instead -- the stage reads phi_stage's stored ghost cells
 * and writes phi_out's face ghost layers: periodic images (code 5) or the
 * extrapolation (codes 1-3) the face fill stores for the face layers (edge
 * and corner ghost cells, which no stage reads, are left as they are), so a
 * sequence of stages needs one lsm3d_amd_fill_ghosts before the first.  Bits
 * 0-2 are ignored.  Where the tiling cannot write them (the last tile along x
 * or y, or the last z-chunk, one cell wide along an extrapolated axis, or
 * one-plane z-chunks), the stage fills phi_out's ghost box after it instead. */
#define LSM_AMD_WRITE_FACE_GHOSTS 8
#define LSM_AMD_RK1 1
#define LSM_AMD_TVDRK2_STAGE2 22
#define LSM_AMD_TVDRK3_STAGE2 32
#define LSM_AMD_TVDRK3_STAGE3 33

/* One stage of phi_t + V.grad(phi) = 0: rhs = -V.grad(phi_stage) with upwind
 * HJ-WENO5, fused with the Runge-Kutta combination.  Equivalent to
 * LSM3D_UPWIND_HJ_WENO5 + LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS +
 * LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS + LSM3D_TVD_RK*_STAGE* in one pass. */
void lsm3d_amd_advect_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                            const LSMLIB_REAL *phi_stage, const int stage_gb[6],
                            const LSMLIB_REAL *phi_cur, const int cur_gb[6],
                            const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                            const LSMLIB_REAL *vel_z, const int vel_gb[6], const int fb[6],
                            LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt,
                            const int bc[6], int periodic_mask);
/* As lsm3d_amd_advect_stage with a constant velocity (per-patch velocity). */
void lsm3d_amd_advect_const_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                                  const LSMLIB_REAL *phi_stage, const int stage_gb[6],
                                  const LSMLIB_REAL *phi_cur, const int cur_gb[6],
                                  LSMLIB_REAL vel_x, LSMLIB_REAL vel_y, LSMLIB_REAL vel_z,
                                  const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                  LSMLIB_REAL dt, const int bc[6], int periodic_mask);
/* One stage of the reinitialization equation phi_t = S(phi)(1 - |grad phi|):
 * HJ-WENO5 plus/minus derivatives of phi_stage, the Godunov right-hand side of
 * LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS (smoothed sign from phi0 when
 * use_phi0_for_sgn == 1, else from phi_stage) and the Runge-Kutta form. */
void lsm3d_amd_reinit_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                            const LSMLIB_REAL *phi_stage, const int stage_gb[6],
                            const LSMLIB_REAL *phi_cur, const int cur_gb[6],
                            const LSMLIB_REAL *phi0, const int phi0_gb[6], int use_phi0_for_sgn,
                            const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                            LSMLIB_REAL dt, const int bc[6], int periodic_mask);
/* LSM3D_COMPUTE_STABLE_ADVECTION_DT that allocates nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and result as LSM3D_COMPUTE_STABLE_ADVECTION_DT
 * (lsm_utilities3d.h, dt = cfl / (max over ib of |u|/dx + |v|/dy + |w|/dz,
 * at least -1, + tiny)), passed by value, except that dt and the velocity
 * arrays must be device memory and dt is written on the device, in stream
 * order.  The partial maxima go to a fixed device array of 65536 doubles
 * (512 KiB per GPU, kept until the process ends); boxes with
 * ceil(nx/64) * ceil(ny/4) > 65536 are rejected with an error.  Every call on
 * one GPU uses that same array: calls that overlap on one GPU (two streams or
 * two host threads at once) race. */
void lsm3d_amd_compute_stable_advection_dt_noalloc(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x,
                                                   const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
                                                   const int vel_gb[6], const int ib[6], LSMLIB_REAL dx,
                                                   LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL cfl);
/* --- Lax-Friedrichs Hamiltonians (no counterpart in the reference, whose
 * level-set RHS routines are upwind or Godunov only) ----------------------
 * From the plus/minus one-sided derivatives (e.g. of LSM3D_HJ_WENO5) and a
 * dissipation coefficient alpha per axis:
 *   advection:    lse_rhs -= sum_a (0.5 (v_a - alpha_a) phi_a+ + 0.5 (v_a + alpha_a) phi_a-)
 *   normal speed: lse_rhs -= vel_n |(grad+ + grad-)/2| - sum_a alpha (phi_a+ - phi_a-)/2
 * alpha_mode 0: local Lax-Friedrichs, alpha_a = |v_a| (|vel_n|) at each cell;
 * with |v_a| >= LSMLIB_ZERO_TOL this is exactly the upwind term of
 * LSM3D_UPWIND_HJ_* + LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS.  alpha_mode 1:
 * global Lax-Friedrichs with the caller's alpha[3] (alpha), which must bound
 * |v_a| (|vel_n|) over the box for stability.  vel_n NULL: the constant
 * vel_n_const.  Work runs over fb; the arrays' ghost boxes must contain it. */
void lsm3d_amd_add_lf_advection_term(LSMLIB_REAL *lse_rhs, const int rhs_gb[6], const LSMLIB_REAL *phi_x_plus,
                                     const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,
                                     const int plus_gb[6], const LSMLIB_REAL *phi_x_minus,
                                     const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus,
                                     const int minus_gb[6], const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                     const LSMLIB_REAL *vel_z, const int vel_gb[6], const int fb[6],
                                     int alpha_mode, const LSMLIB_REAL *alpha);
void lsm3d_amd_add_lf_normal_vel_term(LSMLIB_REAL *lse_rhs, const int rhs_gb[6], const LSMLIB_REAL *phi_x_plus,
                                      const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,
                                      const int plus_gb[6], const LSMLIB_REAL *phi_x_minus,
                                      const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus,
                                      const int minus_gb[6], const LSMLIB_REAL *vel_n, const int vel_n_gb[6],
                                      LSMLIB_REAL vel_n_const, const int fb[6], int alpha_mode,
                                      LSMLIB_REAL alpha);
/* One stage of phi_t + V.grad(phi) = 0 with the Lax-Friedrichs advection
 * Hamiltonian above on the plus/minus HJ-WENO5 derivatives of phi_stage,
 * fused with the Runge-Kutta form (arguments as lsm3d_amd_advect_stage).
 * Equivalent to LSM3D_HJ_WENO5 + LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS +
 * lsm3d_amd_add_lf_advection_term + LSM3D_TVD_RK*_STAGE* in one pass. */
void lsm3d_amd_advect_lf_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                               const LSMLIB_REAL *phi_stage, const int stage_gb[6], const LSMLIB_REAL *phi_cur,
                               const int cur_gb[6], const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                               const LSMLIB_REAL *vel_z, const int vel_gb[6], const int fb[6], LSMLIB_REAL dx,
                               LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6], int periodic_mask,
                               int alpha_mode, const LSMLIB_REAL *alpha);
/* Fills the ghost cells of phi outside fb, face by face in the order x_lo,
 * x_hi, y_lo, y_hi, z_lo, z_hi (each face over the full ghost box in the
 * other directions, as the LSM3D_*_EXTRAPOLATION routines do).  bc[face]:
 * 0 none, 1 homogeneous Neumann (copy), 2 linear, 3 signed linear
 * extrapolation (codes of BoundaryConditionModule.h:125-130), 5 periodic. */
void lsm3d_amd_fill_ghosts(LSMLIB_REAL *phi, const int gb[6], const int fb[6], const int bc[6]);
/* Workgroup tiling of the fused stages: configuration cfg (0 ..
 * lsm3d_amd_stage_tiling_count()-1, shape from lsm3d_amd_stage_tiling_shape)
 * and kz planes per workgroup, for this host thread. */
void lsm3d_amd_set_stage_tiling(int cfg, int kz);
int lsm3d_amd_stage_tiling_count(void);
/* Threshold of the upwind switch (|V| < tol: no advection term) and of the
 * reinitialization cut-off (|phi| < tol: zero right-hand side) used by the
 * fused and patch-list stages of the calling host thread.  The default (and
 * tol < 0) is the build's LSMLIB_ZERO_TOL (1e-11 double, 1e-5 float,
 * configure.ac:264-272), as in the toolbox routines; the float library
 * with 1e-11 follows the double build's switch in FP32 arithmetic. */
void lsm3d_amd_set_zero_tol(double tol);
double lsm3d_amd_zero_tol(void);
void lsm3d_amd_stage_tiling_shape(int cfg, int *tx, int *ty);

/* --- patch lists: a level of uneven patches advanced together ---------
 * Patch q has fill box fbs[6q..6q+5] and ghost box gbs[6q..6q+5] (at least 3
 * ghost cells per side), stored at element offset offsets[q] of every field
 * pool (pool_elems elements each), and the constant velocity
 * vel[3q..3q+2].  The plan keeps device copies of the descriptors; one
 * persistent launch per Runge-Kutta stage covers all patches (the reference
 * loops over patches, LevelSetMethodToolbox.cc:2446-2592).  Ghost values
 * follow bc[6] as for the fused stages: synthesised per face (codes 1 copy,
 * 2 linear, 3 signed linear) or read as stored (code 0); not periodic.
 * Workgroups take tiles from a per-plan device queue that each launch leaves
 * reset, so the launches of one plan must not overlap (one stream, or
 * ordered streams). */
typedef struct lsm3d_amd_patch_plan lsm3d_amd_patch_plan;
lsm3d_amd_patch_plan *lsm3d_amd_patch_plan_create(int npatch, const int *fbs, const int *gbs,
                                                  const long long *offsets, long long pool_elems,
                                                  const LSMLIB_REAL *vel, int tiling_cfg, int kz);
void lsm3d_amd_patch_plan_destroy(lsm3d_amd_patch_plan *plan);
long long lsm3d_amd_patch_plan_cells(const lsm3d_amd_patch_plan *plan);
long long lsm3d_amd_patch_plan_tiles(const lsm3d_amd_patch_plan *plan);
int lsm3d_amd_patch_plan_workgroups(const lsm3d_amd_patch_plan *plan);
/* Writes the face ghost layers of every patch of pool by bc[6] (codes 1-3;
 * 0 leaves a face), for callers that want the extrapolated ghosts stored. */
void lsm3d_amd_patch_plan_fill_ghosts(lsm3d_amd_patch_plan *plan, LSMLIB_REAL *pool, const int bc[6]);
/* One Runge-Kutta stage of phi_t + V.grad(phi) = 0 on every patch. */
void lsm3d_amd_patch_plan_stage(lsm3d_amd_patch_plan *plan, int rk_form, LSMLIB_REAL *out_pool,
                                const LSMLIB_REAL *stage_pool, const LSMLIB_REAL *cur_pool,
                                LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt,
                                const int bc[6]);
/* Captures nsteps TVD-RK3 steps (phi advanced in place through s1, s2)
 * into a hipGraph held by the plan; replay launches it on the library
 * stream.  The graph holds the three pools' device addresses: they must stay
 * allocated while the plan may replay it (until the next capture or the
 * plan's destruction). */
void lsm3d_amd_patch_plan_capture_rk3(lsm3d_amd_patch_plan *plan, int nsteps, LSMLIB_REAL *phi_pool,
                                      LSMLIB_REAL *s1_pool, LSMLIB_REAL *s2_pool, LSMLIB_REAL dx,
                                      LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt,
                                      const int bc[6]);
void lsm3d_amd_patch_plan_replay(lsm3d_amd_patch_plan *plan);

/* --- field extension: S_t + sgn(phi) n . grad S = 0 ---
 * V = signed normal where LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS is active
 * (|phi| > 3 max(dx), or phi of one sign on the six neighbours), else 0, over
 * fb (phi read at fb +- 1).  The fused advection stages with this V advance
 * the extension equation: -V.grad(S), upwind by V, is that routine's
 * right-hand side cell by cell (FieldExtensionAlgorithm.cc:333-440). */
void lsm3d_amd_field_extension_velocity(LSMLIB_REAL *vel_x, LSMLIB_REAL *vel_y, LSMLIB_REAL *vel_z,
                                        const int vel_gb[6], const LSMLIB_REAL *normal_x,
                                        const LSMLIB_REAL *normal_y, const LSMLIB_REAL *normal_z,
                                        const int normal_gb[6], const LSMLIB_REAL *phi, const int phi_gb[6],
                                        const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz);

/* --- curvature model: phi_t + vel_n |grad phi| = b kappa |grad phi|, phi <= mask ---
 * The "medium accuracy" path of LSMLIB's curvature example
 * (examples/serial/curvature_example/curvature_model3d.c:41-626): HJ-ENO2
 * upwinding, TVD-RK2, second-order central curvature, signed linear ghost
 * fill on all six faces after every stage, then the mask.
 * rk_form is 1 (stage 1) or 22 (stage 2; phi_out may alias phi_cur).
 * phi_stage needs valid ghost cells 2 cells out.  With a mask (not NULL),
 * phi_out = max(mask, RK result) on the fill box (IMPOSE_MASK) and phi_raw
 * receives the unmasked result on the two outermost fill-box layers, which
 * lsm3d_amd_extrapolate_ghosts_masked then extrapolates (the reference fills
 * the ghost cells before masking).  zero_tol: lsm3d_amd_set_zero_tol.
 *
 * One stage of the curvature model: LSM3D_HJ_ENO2 +
 * LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS (vel_n; none when |vel_n| <
 * zero_tol) + 4 x LSM3D_CENTRAL_GRAD_ORDER2 + LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS
 * (b; none when 0) + LSM3D_TVD_RK2_STAGE* in one pass (curvature_model3d.c:146-414). */
void lsm3d_amd_curvature_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                               LSMLIB_REAL *phi_raw, const int raw_gb[6], const LSMLIB_REAL *phi_stage,
                               const int stage_gb[6], const LSMLIB_REAL *phi_cur, const int cur_gb[6],
                               const LSMLIB_REAL *mask, const int mask_gb[6], const int fb[6],
                               LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL vel_n,
                               LSMLIB_REAL b, LSMLIB_REAL dt);
/* One stage of the reinitialization equation with HJ-ENO2 (LSM3D_HJ_ENO2 +
 * LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS + LSM3D_TVD_RK2_STAGE*,
 * curvature_model3d.c:484-625), with the optional mask as above. */
void lsm3d_amd_reinit_eno2_stage(int rk_form, LSMLIB_REAL *phi_out, const int out_gb[6],
                                 LSMLIB_REAL *phi_raw, const int raw_gb[6], const LSMLIB_REAL *phi_stage,
                                 const int stage_gb[6], const LSMLIB_REAL *phi_cur, const int cur_gb[6],
                                 const LSMLIB_REAL *phi0, const int phi0_gb[6], int use_phi0_for_sgn,
                                 const LSMLIB_REAL *mask, const int mask_gb[6], const int fb[6],
                                 LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt);
/* The ghost cells of phi (gb outside fb) as the serial package's six-face
 * fill leaves them (signedLinearExtrapolationBC(phi, grid, 9),
 * src/serial/lsm_boundary_conditions.c:306-575, for bc = {3,3,3,3,3,3};
 * codes 1 copy and 2 linear per face as in lsm3d_amd_fill_ghosts), with the
 * fill-box values read from src (NULL: phi itself), then max(mask, value)
 * when mask is not NULL.  Writes ghost cells only. */
void lsm3d_amd_extrapolate_ghosts_masked(LSMLIB_REAL *phi, const int gb[6], const LSMLIB_REAL *src,
                                         const int src_gb[6], const LSMLIB_REAL *mask,
                                         const int mask_gb[6], const int fb[6], const int bc[6]);
/* *dt = LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT of the HJ-ENO2 derivatives
 * of phi (lsm_utilities3d.f.in:455-565) without storing them; phi needs
 * ghost cells 2 cells out.  Synchronizes. */
void lsm3d_amd_const_normal_vel_dt_eno2(LSMLIB_REAL *dt, const LSMLIB_REAL *phi, const int phi_gb[6],
                                        const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                        LSMLIB_REAL vel_n, LSMLIB_REAL cfl);
/* IMPOSE_MASK(phi_masked, mask, phi, grid) of the serial package
 * (src/serial/lsm_macros.h:84-91): phi_masked[q] = max(mask[q], phi[q]) for
 * q < num_gridpts (phi_masked may alias phi). */
void lsm3d_amd_impose_mask(LSMLIB_REAL *phi_masked, const LSMLIB_REAL *mask, const LSMLIB_REAL *phi,
                           long long num_gridpts);

/* --- the narrow-band curvature example, one call per step ------------------
 * curvatureModelMedium3dLocalMainLoop / reinitializeMedium3dLocal
 * (examples/serial/curvature_example/curvature_model3d_local.c:52-728,
 * 757-934) sequenced natively from the narrow-band routines above, on device
 * arrays (every pointer below device memory over gb; the index arrays hold
 * num_gridpts entries; mask may be NULL).  The struct carries the example's
 * state between calls: the level limits and outer-layer limits of the last
 * band, written by the step.  The work arrays keep their contents between
 * calls as the reference's do (the stages read phi_stage1 / phi_next outside
 * the band). */
typedef struct LsmLocalCurvatureModel {
  LSMLIB_REAL *phi_x_plus, *phi_y_plus, *phi_z_plus, *phi_x_minus, *phi_y_minus, *phi_z_minus;
  LSMLIB_REAL *D1, *D2;
  LSMLIB_REAL *phi_x, *phi_y, *phi_z, *phi_xx, *phi_xy, *phi_xz, *phi_yy, *phi_yz, *phi_zz;
  LSMLIB_REAL *lse_rhs, *phi_stage1, *phi_next, *phi0;
  const LSMLIB_REAL *mask;
  unsigned char *narrow_band;
  int *index_x, *index_y, *index_z, *index_outer;
  int gb[6], fb[6];
  double dx[3];
  double a, b, cfl; /* normal speed, curvature coefficient, CFL number */
  int n_lo[6], n_hi[6];
  int outer_limits[4]; /* nlo / nhi of the plus and of the minus outer layer */
} LsmLocalCurvatureModel;
/* One localized TVD-RK2 step of phi (in place, ghost cells included) at
 * interval position dt_sub of tplot; returns dt (-1 on error).  The band is
 * built with the routines above; each stage then runs as one pass over the
 * level-0 entries that recomputes, per point, what the reference's chain of
 * *LOCAL routines (curvature_model3d_local.c:231-580) leaves in its work
 * arrays and reads back, so phi, phi_stage1 and phi_next equal the chain's bit
 * for bit; dt is chosen on the device and read back once.  The derivative
 * work arrays (phi_*_plus/minus, D1, D2, phi_x .. phi_zz) are not written by
 * the step; lse_rhs holds stage 1's right-hand side on level 0. */
double lsm3d_amd_local_curvature_step(LsmLocalCurvatureModel *m, LSMLIB_REAL *phi, double dt_sub, double tplot);
/* The reinitialization trigger after a step (:592-652): 0 none, 1 a sign
 * change in the outer layer, 2 the average |grad phi| outside [0.9, 1.1],
 * 3 TPLOT since the last reinitialization; -1 on error. */
int lsm3d_amd_local_curvature_reinit_trigger(LsmLocalCurvatureModel *m, const LSMLIB_REAL *phi, double t,
                                             double dt_sub, double last_reinit_time);
/* reinitializeMedium3dLocal over levels 0+1 for tmax_r; returns the
 * iteration count (-1 on error). */
int lsm3d_amd_local_curvature_reinit(LsmLocalCurvatureModel *m, LSMLIB_REAL *phi, double tmax_r);

/* ======================================================================
 * Part 4 - reverse mode (no counterpart in the reference, which has no
 * adjoint; these serve gradient-based callers of the WENO5 / TVD-RK
 * advection path).  Device memory only.
 * ====================================================================== */

/* Elements of the work array the reverse stage needs for fill box fb
 * (18 per cell). */
long long lsm3d_amd_advect_adjoint_work_elems(const int fb[6]);

/* Reverse of lsm3d_amd_advect_stage(rk_form, ...) on a fill box periodic on
 * every axis (periodic_mask must be 7; stored ghost cells are not read).
 * With out = a_c phi_cur + a_s (phi_stage + dt r(phi_stage, V)) and adj_out
 * the adjoint of out:
 *   adj_stage (=|+=) a_s (adj_out + dt (dr/dphi_stage)^T adj_out)
 *   adj_cur   (=|+=) a_c adj_out                    (may be NULL for RK1)
 *   vel_adj_{x,y,z} += -a_s dt adj_out D_{x,y,z}    (may all be NULL)
 * accumulate bit 0 selects += for adj_stage, bit 1 for adj_cur.  The upwind
 * stencil choice is taken from the sign of V as in the forward stage. */
void lsm3d_amd_advect_stage_adjoint(int rk_form, const LSMLIB_REAL *adj_out, const int adj_out_gb[6],
                                    LSMLIB_REAL *adj_stage, const int adj_stage_gb[6],
                                    LSMLIB_REAL *adj_cur, const int adj_cur_gb[6], int accumulate,
                                    const LSMLIB_REAL *phi_stage, const int stage_gb[6],
                                    const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                    const LSMLIB_REAL *vel_z, const int vel_gb[6],
                                    LSMLIB_REAL *vel_adj_x, LSMLIB_REAL *vel_adj_y,
                                    LSMLIB_REAL *vel_adj_z, const int vel_adj_gb[6], LSMLIB_REAL *work,
                                    const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                    LSMLIB_REAL dt, int periodic_mask);

/* loss[0] = sum over fb of (H(phi) - H(target))^2 (FP64 accumulation) with
 * the smoothed Heaviside of lsmHeaviside (lsm_calculus_toolbox.f:25-54) of
 * half-width epsilon; grad (optional) receives d loss / d phi. */
void lsm3d_amd_heaviside_mismatch(const LSMLIB_REAL *phi, const int phi_gb[6],
                                  const LSMLIB_REAL *target, const int target_gb[6], const int fb[6],
                                  LSMLIB_REAL epsilon, double *loss, LSMLIB_REAL *grad,
                                  const int grad_gb[6]);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif

#endif /* LSM3D_AMD_H */
