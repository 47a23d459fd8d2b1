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
// dt is chosen on the device (the CFL reduction and the clamps of
// curvature_model3d_local.c:331-349, 383-394, in double precision), so a step
// synchronises with the host once, when the caller reads dt.
#pragma once

#include "lsm_common.h"

namespace lsm {

struct LocalCurvStep {
  const real_t* phi;   // u_n: stage-1 input, u_cur of stage 2
  real_t* phi_stage1;  // stage-1 result, stage-2 input
  real_t* phi_next;    // stage-2 result
  real_t* lse_rhs;     // stage-1 right-hand side on level 0
  const unsigned char* narrow_band;
  const int* index_x;  // band lists, entry l at element l
  const int* index_y;
  const int* index_z;
  int lo0, hi0;  // level-0 entries
  Box gb;        // every array's box
  int mark_fb, mark_d1, mark_d2;
  double a, b, cfl;  // normal speed, curvature coefficient, CFL number
  double dx[3];
  double dt_min_to_correct, dt_min;  // curvature_model3d_local.c:35-36
};

// Stage-1 right-hand side on level 0 and the step's dt (kept on the device).
void local_curv_stage1(const LocalCurvStep& s, double dt_sub, double tplot);
// phi_stage1 = phi + dt lse_rhs on level 0 (LSM3D_TVD_RK2_STAGE1_LOCAL).
void local_curv_rk1(const LocalCurvStep& s);
// Stage-2 right-hand side and phi_next on level 0 (LSM3D_TVD_RK2_STAGE2_LOCAL).
void local_curv_stage2(const LocalCurvStep& s);
// The step's dt, read back after the stream has drained; -1 on error.
double local_curv_dt(const LocalCurvStep& s);

}  // namespace lsm
