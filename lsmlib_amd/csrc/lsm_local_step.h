// The two TVD-RK2 stages of the narrow-band curvature model's step
// (curvature_model3d_local.c:231-580), each as one pass over the level-0
// entries of the band instead of the reference's chain of *LOCAL routines.
//
// Per stage the reference zeroes lse_rhs, runs HJ_ENO2_LOCAL (D1 on levels
// 0-2, D2 on levels 0-1, the one-sided derivatives on level 0), the constant
// normal-speed term, (stage 1) the normal-speed CFL dt, CENTRAL_GRAD_ORDER2
// on levels 0-1 and on its own output on level 0, the curvature term, the
// cut-off and the RK2 stage on level 0.  Every value a level-0 point reads
// from those work arrays lies within two cells along an axis (or one
// diagonal cell) of the point, so on a band built by
// LSM3D_DETERMINE_NARROW_BAND with the D2 / D1 boundary-layer marks of the
// model each is written by the same call from phi alone; the fused pass
// recomputes them per point with the same operations in the same order
// (1e10 where the mark excludes a difference, as lsm3dComputeDnLOCAL writes),
// so phi_stage1 and phi_next equal the routine chain's bit for bit.  A point
// that would need a value the chain leaves unwritten (a neighbour outside
// the marks the model sets) is counted as skipped, as the routines count
// entries outside their arrays.
//
// The band's level sizes stay on the device while it is built (kernels sized
// by them run on fixed grids and stride over their work) and dt is chosen on
// the device (the CFL reduction and the clamps of
// curvature_model3d_local.c:331-349, 383-394, in double precision), so a step
// synchronises with the host once, when the limits and dt are read back.
#pragma once

#include "lsm_common.h"

namespace lsm {

struct LocalCurvStep {
  const real_t* phi;   // u_n: stage-1 input, u_cur of stage 2
  real_t* phi_stage1;  // stage-1 result, stage-2 input
  real_t* phi_next;    // stage-2 result
  real_t* lse_rhs;     // stage-1 right-hand side on level 0
  const real_t* mask;  // NULL: no mask
  unsigned char* narrow_band;
  int* index_x;  // band lists, entry l at element l (0 .. num_points - 1)
  int* index_y;
  int* index_z;
  int* index_outer;  // outer-layer lists, entries 0 .. num_points - 1
  int num_points;
  Box gb;  // every array's box
  int mark_fb, mark_d1, mark_d2;
  int levels;            // band levels 0 .. levels
  int prev_lo[8], prev_hi[8];  // level limits kept when level 0 is empty
  double a, b, cfl;  // normal speed, curvature coefficient, CFL number
  double dx[3];
  double dt_min_to_correct, dt_min;  // curvature_model3d_local.c:35-36
};

// What a step leaves on the host: the band's level limits and outer-layer
// sizes and dt.
struct LocalStepResult {
  int n_lo[8], n_hi[8];
  int tot[3];  // level 0: band points, outer minus, outer plus
  double dt;
};

// LSM3D_DETERMINE_NARROW_BAND (width gamma = 4 dx, inner width beta = 2 dx,
// `levels` levels) with the level sizes kept on the device: the same lists,
// marks and limits, built without synchronising with the host.
void local_band(const LocalCurvStep& s);
// Stage-1 right-hand side on level 0 and the step's dt (kept on the device).
void local_curv_stage1(const LocalCurvStep& s, double dt_sub, double tplot);
// phi_stage1 = phi + dt lse_rhs on level 0 (LSM3D_TVD_RK2_STAGE1_LOCAL).
void local_curv_rk1(const LocalCurvStep& s);
// Stage-2 right-hand side and phi_next on level 0 (LSM3D_TVD_RK2_STAGE2_LOCAL).
void local_curv_stage2(const LocalCurvStep& s);
// IMPOSE_MASK_LOCAL over level-0 entries n_lo[0] .. n_hi[0] - 1: dest = max(mask, src).
void local_mask(const LocalCurvStep& s, real_t* dest, const real_t* src);
// Reads the band limits and dt back once the stream has drained; false on
// error (a list too short, or a pending error).
bool local_step_result(const LocalCurvStep& s, LocalStepResult* r);
// The reinitialization trigger's two reductions (LSM3D_CHECK_OUTER_NARROW_BAND_LAYER
// and LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL over level 0 = entries nlo0 .. nhi0),
// with their kernels and launch shapes, read back with one synchronisation.
bool local_trigger_values(const LocalCurvStep& s, int nlo0, int nhi0, const int outer_limits[4], int* change,
                          real_t* ave);

}  // namespace lsm
