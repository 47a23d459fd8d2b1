// The narrow-band curvature example's step sequence as native host code.
//
// curvatureModelMedium3dLocalMainLoop and reinitializeMedium3dLocal
// (examples/serial/curvature_example/curvature_model3d_local.c:52-728,
// 757-934) call some forty toolbox routines per step with host-side
// decisions between them (dt, the level limits, the reinitialization
// trigger).  Sequencing them here, on device arrays and the library stream,
// leaves the caller one call per step; every routine called is the drop-in
// entry point of the same name, so results equal the reference's sequence of
// those routines.
#include <cmath>
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_local_step.h"
#include "lsm_runtime.h"

namespace {

typedef real_t R;
typedef LsmLocalCurvatureModel Model;

constexpr unsigned char kMarkGb = 127, kMarkD1 = 126, kMarkD2 = 125, kMarkFb = 124;  // :43
constexpr int kLevel = 3;
constexpr double kDtMinToCorrect = 1e-5, kDtMin = 0.001;   // curvature_model3d_local.c:35-36
constexpr double kAveGradMin = 0.9, kAveGradMax = 1.1;      // :41-42
constexpr double kTplot = 0.1;                              // curvature_model_top.h:15

#define GB6(m) &(m)->gb[0], &(m)->gb[1], &(m)->gb[2], &(m)->gb[3], &(m)->gb[4], &(m)->gb[5]
#define B6(b) &(b)[0], &(b)[1], &(b)[2], &(b)[3], &(b)[4], &(b)[5]
#define FAIL_IF_ERROR(ret) \
  if (lsm::error_pending()) return ret

struct Scalars {  // the by-reference scalars of the routines, in the library's precision
  R dx, dy, dz, a, b, cfl, beta, gamma;
};

Scalars scalars(const Model* m) {
  return Scalars{(R)m->dx[0], (R)m->dx[1], (R)m->dx[2], (R)m->a, (R)m->b, (R)m->cfl, (R)(2.0 * m->dx[0]),
                 (R)(4.0 * m->dx[0])};
}

int num_points(const Model* m) {
  return (m->gb[1] - m->gb[0] + 1) * (m->gb[3] - m->gb[2] + 1) * (m->gb[5] - m->gb[4] + 1);
}

void ghost_fill(Model* m, R* a) {
  static const int bc[6] = {3, 3, 3, 3, 3, 3};
  lsm3d_amd_extrapolate_ghosts_masked(a, m->gb, nullptr, m->gb, nullptr, m->gb, m->fb, bc);
}

// IMPOSE_MASK_LOCAL (src/serial/lsm_macros.h:108-117): entries n_lo[0] .. n_hi[0]-1
void mask_local(Model* m, R* dest, const R* src) {
  int lo = m->n_lo[0], hi = m->n_hi[0] - 1;
  // the reference prototype takes src and mask non-const; the routine only reads them
  lsm3dimposemasklocal_(dest, GB6(m), const_cast<R*>(src), GB6(m), const_cast<R*>(m->mask), GB6(m), m->index_x,
                        m->index_y, m->index_z, &lo, &hi);
}

void eno2(Model* m, const R* phi, const Scalars& s) {
  lsm3dhjeno2local_(m->phi_x_plus, m->phi_y_plus, m->phi_z_plus, GB6(m), m->phi_x_minus, m->phi_y_minus,
                    m->phi_z_minus, GB6(m), phi, GB6(m), m->D1, GB6(m), m->D2, GB6(m), &s.dx, &s.dy, &s.dz,
                    m->index_x, m->index_y, m->index_z, &m->n_lo[0], &m->n_hi[0], &m->n_lo[1], &m->n_hi[1],
                    &m->n_lo[2], &m->n_hi[2], m->narrow_band, GB6(m), &kMarkFb, &kMarkD1, &kMarkD2);
}

// The fused step's view of the model: its arrays, the marks of band(), the
// model's coefficients, and the level limits a band build with an empty
// level 0 leaves as they are.
lsm::LocalCurvStep fused_step(const Model* m, const R* phi) {
  lsm::LocalCurvStep st;
  st.phi = phi;
  st.phi_stage1 = m->phi_stage1;
  st.phi_next = m->phi_next;
  st.lse_rhs = m->lse_rhs;
  st.mask = m->mask;
  st.narrow_band = m->narrow_band;
  st.index_x = m->index_x;
  st.index_y = m->index_y;
  st.index_z = m->index_z;
  st.index_outer = m->index_outer;
  st.num_points = num_points(m);
  st.gb = lsm::Box{m->gb[0], m->gb[1], m->gb[2], m->gb[3], m->gb[4], m->gb[5]};
  st.mark_fb = kMarkFb;
  st.mark_d1 = kMarkD1;
  st.mark_d2 = kMarkD2;
  st.levels = kLevel;
  for (int l = 0; l < 8; l++) {
    st.prev_lo[l] = l < 6 ? m->n_lo[l] : 0;
    st.prev_hi[l] = l < 6 ? m->n_hi[l] : 0;
  }
  st.a = m->a;
  st.b = m->b;
  st.cfl = m->cfl;
  for (int q = 0; q < 3; q++) st.dx[q] = m->dx[q];
  st.dt_min_to_correct = kDtMinToCorrect;
  st.dt_min = kDtMin;
  return st;
}

// The boundary-layer marks of band() (no host synchronisation).
void band_marks(Model* m) {
  int d2[6], d1[6];
  for (int q = 0; q < 6; q++) {
    d2[q] = m->gb[q] + ((q & 1) ? -2 : 2);
    d1[q] = m->gb[q] + ((q & 1) ? -1 : 1);
  }
  lsm3dmarknarrowbandboundarylayer_(m->narrow_band, GB6(m), B6(d2), &kMarkD2);
  FAIL_IF_ERROR();
  lsm3dmarknarrowbandboundarylayer_(m->narrow_band, GB6(m), B6(d1), &kMarkD1);
  FAIL_IF_ERROR();
  lsm3dmarknarrowbandboundarylayer_(m->narrow_band, GB6(m), GB6(m), &kMarkGb);
}

// The model's arrays must be device memory (the calls below take them as such
// without asking the runtime).  Checked when the set of pointers differs from
// the last set this thread validated.
bool model_on_device(const Model* m, const void* phi) {
  const void* ptrs[] = {m->phi_x_plus, m->phi_y_plus, m->phi_z_plus, m->phi_x_minus, m->phi_y_minus,
                        m->phi_z_minus, m->D1, m->D2, m->phi_x, m->phi_y, m->phi_z, m->phi_xx, m->phi_xy,
                        m->phi_xz, m->phi_yy, m->phi_yz, m->phi_zz, m->lse_rhs, m->phi_stage1, m->phi_next,
                        m->phi0, m->narrow_band, m->index_x, m->index_y, m->index_z, m->index_outer, phi,
                        m->mask};
  constexpr int n = sizeof(ptrs) / sizeof(ptrs[0]);
  static thread_local uint64_t validated = 0;
  uint64_t sig = 1469598103934665603ull;  // FNV-1a over the pointer values
  for (int q = 0; q < n; q++) sig = (sig ^ (uint64_t)(uintptr_t)ptrs[q]) * 1099511628211ull;
  if (sig == validated) return true;
  for (int q = 0; q < n; q++) {
    if (q == n - 1 && ptrs[q] == nullptr) continue;  // no mask
    if (ptrs[q] == nullptr || !lsm::is_device_pointer(ptrs[q])) {
      lsm::raise_error("lsm3d_amd_local_curvature: every array of the model (and phi) must be device memory");
      return false;
    }
  }
  validated = sig;
  return true;
}

}  // namespace

extern "C" {

double lsm3d_amd_local_curvature_step(LsmLocalCurvatureModel* m, LSMLIB_REAL* phi, double dt_sub, double tplot) {
  lsm::clear_error();
  if (!model_on_device(m, phi)) return -1.0;
  lsm::AssumeDevice on_device;
  // band(), its limits kept on the device, then the two stages fused per
  // level-0 point (lsm_local_step.h); one host synchronisation at the end
  const lsm::LocalCurvStep st = fused_step(m, phi);
  lsm::local_band(st);
  FAIL_IF_ERROR(-1.0);
  band_marks(m);
  FAIL_IF_ERROR(-1.0);
  lsm::local_curv_stage1(st, dt_sub, tplot);
  FAIL_IF_ERROR(-1.0);
  lsm::local_curv_rk1(st);
  FAIL_IF_ERROR(-1.0);
  ghost_fill(m, m->phi_stage1);
  FAIL_IF_ERROR(-1.0);
  if (m->mask) {
    lsm::local_mask(st, m->phi_stage1, m->phi_stage1);
    FAIL_IF_ERROR(-1.0);
  }
  lsm::local_curv_stage2(st);
  FAIL_IF_ERROR(-1.0);
  ghost_fill(m, m->phi_next);
  FAIL_IF_ERROR(-1.0);
  if (m->mask) {
    lsm::local_mask(st, phi, m->phi_next);
  } else {
    const size_t bytes = (size_t)num_points(m) * sizeof(R);
    lsm::check_hip(hipMemcpyAsync(phi, m->phi_next, bytes, hipMemcpyDeviceToDevice, lsm::current_stream()),
                   "hipMemcpyAsync(phi <- phi_next)");
  }
  FAIL_IF_ERROR(-1.0);
  lsm::LocalStepResult r;
  if (!lsm::local_step_result(st, &r)) return -1.0;
  // the limits LSM3D_DETERMINE_NARROW_BAND returns (lsm_localization3d.f:232-298)
  for (int l = 0; l <= kLevel; l++) {
    m->n_lo[l] = r.n_lo[l];
    m->n_hi[l] = r.n_hi[l];
  }
  const int last = num_points(m) - 1;
  int* ol = m->outer_limits;  // nlo_plus, nhi_plus, nlo_minus, nhi_minus
  ol[2] = 0;
  ol[1] = last;
  if (r.tot[0] > 0) {
    ol[3] = r.tot[1] - 1;
    ol[0] = last - r.tot[2] + 1;
  } else {
    ol[3] = 0;
    ol[0] = last;
  }
  return r.dt;
}

int lsm3d_amd_local_curvature_reinit_trigger(LsmLocalCurvatureModel* m, const LSMLIB_REAL* phi, double t,
                                             double dt_sub, double last_reinit_time) {
  lsm::clear_error();
  if (!model_on_device(m, phi)) return -1;
  lsm::AssumeDevice on_device;
  // LSM3D_CHECK_OUTER_NARROW_BAND_LAYER, then LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL
  // (evaluated together, one read-back; the average is used only without a
  // sign change, as in the example)
  int change = 0;
  R ave = 0;
  if (!lsm::local_trigger_values(fused_step(m, phi), m->n_lo[0], m->n_hi[0], m->outer_limits, &change, &ave))
    return -1;
  if (change) return 1;
  if ((double)ave < kAveGradMin || (double)ave > kAveGradMax) return 2;
  return (t + dt_sub - last_reinit_time >= kTplot) ? 3 : 0;
}

int lsm3d_amd_local_curvature_reinit(LsmLocalCurvatureModel* m, LSMLIB_REAL* phi, double tmax_r) {
  lsm::clear_error();
  if (!model_on_device(m, phi)) return -1;
  lsm::AssumeDevice on_device;
  const Scalars s = scalars(m);
  int lo[6], hi[6];
  for (int q = 0; q < 6; q++) {
    lo[q] = m->n_lo[q];
    hi[q] = m->n_hi[q];
  }
  // levels 0 and 1 act as level 0 (curvature_model3d_local.c:661-663)
  m->n_hi[0] = m->n_hi[1];
  m->n_lo[1] = m->n_lo[2];
  m->n_hi[1] = m->n_hi[2];
  m->n_lo[2] = m->n_lo[3];
  m->n_hi[2] = m->n_hi[3];
  const size_t bytes = (size_t)num_points(m) * sizeof(R);
  lsm::check_hip(hipMemcpyAsync(m->phi0, phi, bytes, hipMemcpyDeviceToDevice, lsm::current_stream()),
                 "hipMemcpyAsync(phi0 <- phi)");
  const double dt_r = m->cfl * m->dx[0];
  const R dtr = (R)dt_r;
  const int use_phi0 = 0;
  double t_r = 0.0;
  int n = 0;
  while (t_r < tmax_r && !lsm::error_pending()) {
    for (int stage = 1; stage <= 2 && !lsm::error_pending(); stage++) {
      const R* src = (stage == 1) ? phi : m->phi_stage1;
      eno2(m, src, s);
      if (lsm::error_pending()) break;
      lsm3dcomputereinitializationeqnrhslocal_(m->lse_rhs, GB6(m), src, GB6(m), m->phi0, GB6(m), m->phi_x_plus,
                                               m->phi_y_plus, m->phi_z_plus, GB6(m), m->phi_x_minus,
                                               m->phi_y_minus, m->phi_z_minus, GB6(m), &s.dx, &s.dy, &s.dz,
                                               &use_phi0, m->index_x, m->index_y, m->index_z, &m->n_lo[0],
                                               &m->n_hi[0], m->narrow_band, GB6(m), &kMarkFb);
      if (lsm::error_pending()) break;
      if (stage == 1) {
        lsm3dtvdrk2stage1local_(m->phi_stage1, GB6(m), phi, GB6(m), m->lse_rhs, GB6(m), &dtr, m->index_x,
                                m->index_y, m->index_z, &m->n_lo[0], &m->n_hi[0], m->narrow_band, GB6(m), &kMarkFb);
        if (lsm::error_pending()) break;
        ghost_fill(m, m->phi_stage1);
      } else {
        lsm3dtvdrk2stage2local_(m->phi_next, GB6(m), m->phi_stage1, GB6(m), phi, GB6(m), m->lse_rhs, GB6(m), &dtr,
                                m->index_x, m->index_y, m->index_z, &m->n_lo[0], &m->n_hi[0], m->narrow_band,
                                GB6(m), &kMarkFb);
        if (lsm::error_pending()) break;
        ghost_fill(m, m->phi_next);
        if (lsm::error_pending()) break;
        if (m->mask) {
          mask_local(m, phi, m->phi_next);
        } else {
          lsm::check_hip(hipMemcpyAsync(phi, m->phi_next, bytes, hipMemcpyDeviceToDevice, lsm::current_stream()),
                         "hipMemcpyAsync(phi <- phi_next)");
        }
      }
    }
    t_r = t_r + dt_r;
    n++;
  }
  for (int q = 0; q < 6; q++) {
    m->n_lo[q] = lo[q];
    m->n_hi[q] = hi[q];
  }
  return lsm::error_pending() ? -1 : n;
}

}  // extern "C"
