// Fused TVD Runge-Kutta stages on device-resident fields.
//
// One launch advances one stage over the whole fill box:
//   advection:  phi_out = RK(phi_cur, phi_s, -(u phi_x + v phi_y + w phi_z))
//               phi_{x,y,z} = upwind HJ-WENO5 of phi_s
//               (the reference path LSM3D_UPWIND_HJ_WENO5 ->
//                LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS -> LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS
//                -> LSM3D_TVD_RK3_STAGE*, LevelSetFunctionIntegrator.cc:1319-1510)
//   reinit:     phi_out = RK(phi_cur, phi_s, S(phi)(1 - |grad phi|*))
//               with HJ-WENO5 plus/minus derivatives of phi_s
//               (LSM3D_HJ_WENO5 -> LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS -> RK,
//                ReinitializationAlgorithm.cc:698-1020)
// Each stage reads phi_s, the velocity (or phi0) and phi_cur once from HBM
// and writes phi_out once; undivided differences, one-sided derivatives and
// the right-hand side stay in registers and LDS.  Ghost cells of phi_s are
// either read as stored or synthesised from the fill box per face
// (extrapolation or periodic image), which removes the ghost-fill pass
// between stages.
#include <cstdio>
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_march.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

namespace lsm {

// SYNTH: ghost values are synthesised per face; otherwise stored ghost cells
// are read.  FP32 uses the rearranged WENO5 form, FP64 the reference order.
// FACES (with stored ghosts only): the stage writes its output's
// extrapolated face ghost layers (a.oghost).
template <typename T, int TX, int TY, int KIND, bool SYNTH, bool FACES = false>
__global__ void __launch_bounds__(TX* TY) k_stage(StageArgs<T> a) {
  __shared__ StageTile<T, TX, TY> tile;
  int bx, by, bz;
  launch_tile_coords(a.tg, (int)blockIdx.x, (int)gridDim.x, bx, by, bz);
  if (SYNTH && !tile_is_interior<TX, TY>(a.fb, a.tg, bx, by, bz))
    march_tile<T, TX, TY, KIND, (sizeof(T) == 4), true>(a, bx, by, bz, tile);
  else
    march_tile<T, TX, TY, KIND, (sizeof(T) == 4), false, true, FACES && !SYNTH>(a, bx, by, bz, tile);
}

// ---------------------------------------------------------------------------
// Ghost fill of one axis (both faces) on device.  bc codes per face:
// 0 none, 1 homogeneous Neumann / copy, 2 linear, 3 signed linear
// (lsm_boundary_conditions3d.f:46-640), 5 periodic.  Threads cover the full
// ghost box in the two other directions, like the reference face loops.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fill_axis(Field<T> phi, Box gb, Box fb, int d, int bc_lo, int bc_hi) {
  const int u = (d + 1) % 3, v = (d + 2) % 3;
  const int glo[3] = {gb.ilo, gb.jlo, gb.klo}, ghi[3] = {gb.ihi, gb.jhi, gb.khi};
  const int flo[3] = {fb.ilo, fb.jlo, fb.klo}, fhi[3] = {fb.ihi, fb.jhi, fb.khi};
  const int pu = glo[u] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int pv = glo[v] + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int hi = (int)blockIdx.z;
  if (pu > ghi[u] || pv > ghi[v]) return;
  const int bc = hi ? bc_hi : bc_lo;
  if (bc == 0) return;
  const int n = fhi[d] - flo[d] + 1;
  int idx[3];
  idx[u] = pu;
  idx[v] = pv;
  const int lo = hi ? fhi[d] + 1 : glo[d], up = hi ? ghi[d] : flo[d] - 1;
  if (bc == 5) {
    for (int t = lo; t <= up; t++) {
      idx[d] = hi ? t - n : t + n;
      const T val = phi.at(idx[0], idx[1], idx[2]);
      idx[d] = t;
      phi.at(idx[0], idx[1], idx[2]) = val;
    }
    return;
  }
  const int bnd = hi ? fhi[d] : flo[d];
  idx[d] = bnd;
  const T pb = phi.at(idx[0], idx[1], idx[2]);
  idx[d] = bnd + (hi ? -1 : 1);
  const T p1 = phi.at(idx[0], idx[1], idx[2]);
  for (int t = lo; t <= up; t++) {
    idx[d] = t;
    phi.at(idx[0], idx[1], idx[2]) = extrapolate<T>(bc, pb, p1, hi ? (t - bnd) : (bnd - t));
  }
}

}  // namespace lsm

// ===========================================================================
// Host side
// ===========================================================================
namespace lsm {

// Tile configurations compiled in (TX x TY threads).  The autotuner picks
// one per grid-size class through lsm3d_amd_set_stage_tiling().
constexpr TileShape kTiles[] = {{64, 4}, {64, 8}, {32, 8}, {128, 2}, {128, 4}, {32, 16}, {16, 16}, {16, 8}, {16, 4}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

thread_local Tiling tl_tiling = {0, 32};
// Threshold of the upwind switch (|V| < tol -> no advection term) and of the
// reinitialization cut-off (|phi| < tol -> zero rhs).  The default is the
// build's LSMLIB_ZERO_TOL (configure.ac:264-272), as in the toolbox routines.
thread_local double tl_zero_tol = Tol<real_t>::zero;

Tiling current_tiling() { return tl_tiling; }
double current_zero_tol() { return tl_zero_tol; }
TileShape tile_shape(int cfg) { return kTiles[cfg]; }
int tiling_count() { return kNumTiles; }

bool valid_form(int form) {
  return form == RK_EULER || form == RK2_STAGE2 || form == RK3_STAGE2 || form == RK3_STAGE3;
}

bool device_only(const char* fn, const void* p, const char* what) {
  if (p && !assume_device() && !is_device_pointer(p)) {
    raise_error(std::string(fn) + ": " + what + " must be device memory");
    return false;
  }
  return true;
}

// Validates the 6 ghost codes for a fill box; returns the rule.
GhostRule make_ghost_rule(const char* fn, const int* bc, const Box& fb) {
  GhostRule g{};
  if (!bc) return g;
  const int n[3] = {fb.nx(), fb.ny(), fb.nz()};
  for (int f = 0; f < 6; f++) {
    const int c = bc[f];
    if (c != 0 && c != 1 && c != 2 && c != 3 && c != 5) raise_error(std::string(fn) + ": unknown ghost code");
    if ((c == 2 || c == 3) && n[f / 2] < 2) raise_error(std::string(fn) + ": extrapolation needs 2 cells");
    if (c == 5 && n[f / 2] < 3) raise_error(std::string(fn) + ": a periodic axis needs 3 cells");
    g.code[f] = c;
    g.any |= c;
  }
  for (int d = 0; d < 3; d++)
    if ((bc[2 * d] == 5) != (bc[2 * d + 1] == 5))
      raise_error(std::string(fn) + ": periodic faces come in pairs");
  return g;
}

Field<const real_t> cfield(const real_t* p, const int* gb) {
  Field<const real_t> f;
  if (!p) {
    f.p = nullptr;
    f.ilo = f.jlo = f.klo = 0;
    f.sy = f.sz = 0;
    return f;
  }
  return make_field<const real_t>(p, Box{gb[0], gb[1], gb[2], gb[3], gb[4], gb[5]});
}

namespace {
typedef real_t R;

Box box6(const int* b) { return Box{b[0], b[1], b[2], b[3], b[4], b[5]}; }

template <int TX, int TY, int KIND>
void launch_stage_t(hipStream_t s, const StageArgs<R>& a, int nblk) {
  if (a.ghost.any)
    hipLaunchKernelGGL((k_stage<R, TX, TY, KIND, true>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
  else if (a.oghost.any)
    hipLaunchKernelGGL((k_stage<R, TX, TY, KIND, false, true>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
  else
    hipLaunchKernelGGL((k_stage<R, TX, TY, KIND, false>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
}

template <int KIND>
void launch_stage_kind(hipStream_t s, const StageArgs<R>& a, int cfg, int nblk) {
  switch (cfg) {
    case 0: launch_stage_t<64, 4, KIND>(s, a, nblk); break;
    case 1: launch_stage_t<64, 8, KIND>(s, a, nblk); break;
    case 2: launch_stage_t<32, 8, KIND>(s, a, nblk); break;
    case 3: launch_stage_t<128, 2, KIND>(s, a, nblk); break;
    case 4: launch_stage_t<128, 4, KIND>(s, a, nblk); break;
    case 5: launch_stage_t<32, 16, KIND>(s, a, nblk); break;
    case 6: launch_stage_t<16, 16, KIND>(s, a, nblk); break;
    case 7: launch_stage_t<16, 8, KIND>(s, a, nblk); break;
    default: launch_stage_t<16, 4, KIND>(s, a, nblk); break;
  }
}

// The ghost box of phi outside the fill box, face by face in the order x, y,
// z (the reference's fill order; codes already validated).
void fill_ghost_box(Call& c, R* phi, const int* gb, const int* fb, const int* bc) {
  const Box g = box6(gb), f = box6(fb);
  Field<R> fp = make_field<R>(phi, g);
  const int ext[3] = {g.nx(), g.ny(), g.nz()};
  for (int d = 0; d < 3; d++) {
    if (bc[2 * d] == 0 && bc[2 * d + 1] == 0) continue;
    const int u = (d + 1) % 3, v = (d + 2) % 3;
    dim3 grid((unsigned)((ext[u] + 63) / 64), (unsigned)((ext[v] + 3) / 4), 2);
    hipLaunchKernelGGL(k_fill_axis<R>, grid, dim3(64, 4, 1), 0, c.stream(), fp, g, f, d, bc[2 * d], bc[2 * d + 1]);
    c.launched("k_fill_axis");
  }
}

// Shared by the advection and reinitialization stage entry points.
void run_stage(const char* fn, int kind, int form, R* out, const int* out_gb, const R* ps,
               const int* ps_gb, const R* pc, const int* pc_gb, const R* vx, const R* vy, const R* vz,
               const int* vel_gb, const R* phi0, const int* phi0_gb, R cvx, R cvy, R cvz,
               const int* fbp, R dx, R dy, R dz, R dt, int use_phi0, const int* bc, int periodic,
               int alpha_mode = 0, const R* alpha = nullptr) {
  Call c(fn);
  const Box fb = box6(fbp);
  if (!valid_form(form)) raise_error(std::string(fn) + ": unknown Runge-Kutta form");
  // tiles read phi_stage's halo while neighbouring tiles write phi_out
  if (out && (const void*)out == (const void*)ps) raise_error(std::string(fn) + ": phi_out must not alias phi_stage");
  require_inside(fn, "phi_out", box6(out_gb), fb, 0, 0, 0);
  require_inside(fn, "phi_stage", box6(ps_gb), fb, 3, 3, 3);
  if (pc) require_inside(fn, "phi_cur", box6(pc_gb), fb, 0, 0, 0);
  if (kind == STAGE_ADVECT_FIELD || kind == STAGE_ADVECT_LF) require_inside(fn, "vel", box6(vel_gb), fb, 0, 0, 0);
  if (kind == STAGE_ADVECT_LF) {
    if (alpha_mode != 0 && alpha_mode != 1) raise_error(std::string(fn) + ": alpha_mode is 0 (local) or 1 (global)");
    if (alpha_mode == 1 && !alpha) raise_error(std::string(fn) + ": global alpha needs alpha[3]");
    if (alpha_mode == 1 && alpha && !(alpha[0] >= 0 && alpha[1] >= 0 && alpha[2] >= 0))
      raise_error(std::string(fn) + ": alpha must be non-negative");
  }
  if (kind == STAGE_REINIT && use_phi0) require_inside(fn, "phi0", box6(phi0_gb), fb, 0, 0, 0);
  // bit 8: bc gives phi_out's face ghost codes; phi_stage's stored ghosts are read
  const bool face_out = (periodic & 8) != 0;
  if (face_out && !bc) raise_error(std::string(fn) + ": writing face ghosts (periodic_mask bit 8) needs bc");
  GhostRule ghost = make_ghost_rule(fn, bc, fb);
  GhostRule oghost{};
  bool fill_after = false;  // the tiling cannot write the face ghosts: fill phi_out's ghosts afterwards
  if (face_out && bc) {
    periodic = (bc[0] == 5 ? 1 : 0) | (bc[2] == 5 ? 2 : 0) | (bc[4] == 5 ? 4 : 0);
    const TileShape ts = kTiles[tl_tiling.cfg];
    for (int f = 0; f < 6; f++) oghost.code[f] = (ghost.code[f] == 5) ? 0 : ghost.code[f];
    oghost.any = oghost.code[0] | oghost.code[1] | oghost.code[2] | oghost.code[3] | oghost.code[4] | oghost.code[5];
    // a hi face cell and its inward neighbour must share a tile (x, y) and a
    // z-chunk (z): the last tile or chunk along an extrapolated axis is not one
    // cell wide, and z-chunks have two planes at least
    const int kz = tl_tiling.kz;
    if ((oghost.code[1] && fb.nx() % ts.tx == 1) || (oghost.code[3] && fb.ny() % ts.ty == 1) ||
        ((oghost.code[4] || oghost.code[5]) && (kz < 2 || fb.nz() % kz == 1))) {
      fill_after = oghost.any != 0;
      oghost = GhostRule{};
    }
    ghost = GhostRule{};
    require_inside(fn, "phi_out", box6(out_gb), fb, 3, 3, 3);
  }
  periodic &= 7;
  if (periodic) {
    require_inside(fn, "phi_out", box6(out_gb), fb, 3, 3, 3);
    const int n[3] = {fb.nx(), fb.ny(), fb.nz()};
    for (int d = 0; d < 3; d++)
      if (((periodic >> d) & 1) && n[d] < 6) raise_error(std::string(fn) + ": periodic images need 6 cells");
  }
  if (!(device_only(fn, out, "phi_out") && device_only(fn, ps, "phi_stage") &&
        device_only(fn, pc, "phi_cur") && device_only(fn, vx, "vel_x") && device_only(fn, vy, "vel_y") &&
        device_only(fn, vz, "vel_z") && device_only(fn, phi0, "phi0")))
    return;
  if (!c.ok() || fb.cells() == 0) {
    c.finish();
    return;
  }
  const Tiling tl = tl_tiling;
  const TileShape ts = kTiles[tl.cfg];
  StageArgs<R> a;
  a.out = make_field<R>(out, box6(out_gb));
  a.ps = cfield(ps, ps_gb);
  a.pc = cfield(pc, pc_gb);
  const bool vel_field = kind == STAGE_ADVECT_FIELD || kind == STAGE_ADVECT_LF;
  a.vx = cfield(vel_field ? vx : nullptr, vel_gb);
  a.vy = cfield(vel_field ? vy : nullptr, vel_gb);
  a.vz = cfield(vel_field ? vz : nullptr, vel_gb);
  a.phi0 = cfield((kind == STAGE_REINIT && use_phi0) ? phi0 : nullptr, phi0_gb);
  a.cvx = cvx;
  a.cvy = cvy;
  a.cvz = cvz;
  a.fb = fb;
  a.tg = make_tile_grid(fb, ts.tx, ts.ty, tl.kz);
  a.ghost = ghost;
  a.ix = R(1) / dx;
  a.iy = R(1) / dy;
  a.iz = R(1) / dz;
  a.tdx2[0] = R(12) * Tol<R>::tiny / (a.ix * a.ix);
  a.tdx2[1] = R(12) * Tol<R>::tiny / (a.iy * a.iy);
  a.tdx2[2] = R(12) * Tol<R>::tiny / (a.iz * a.iz);
  a.dt = dt;
  R mdx = dx;
  if (dy > mdx) mdx = dy;
  if (dz > mdx) mdx = dz;
  a.dx_sq = mdx * mdx;  // lsm_reinitialization3d.f.in:140-141
  a.zero_tol = (R)tl_zero_tol;
  a.form = form;
  a.rkb = rk_blend_weight<R>(form);
  a.periodic = periodic;
  a.oghost = oghost;
  a.use_phi0 = use_phi0 ? 1 : 0;
  a.alpha_mode = (kind == STAGE_ADVECT_LF) ? alpha_mode : 0;
  for (int d = 0; d < 3; d++) a.alpha[d] = (a.alpha_mode && alpha) ? alpha[d] : R(0);
  const long long nblk = (long long)a.tg.ntx * a.tg.nty * a.tg.nkz;
  if (nblk > 0x7fffffffLL) {
    raise_error(std::string(fn) + ": grid too large for one launch");
    c.finish();
    return;
  }
  if (kind == STAGE_ADVECT_FIELD) launch_stage_kind<STAGE_ADVECT_FIELD>(c.stream(), a, tl.cfg, (int)nblk);
  else if (kind == STAGE_ADVECT_CONST) launch_stage_kind<STAGE_ADVECT_CONST>(c.stream(), a, tl.cfg, (int)nblk);
  else if (kind == STAGE_ADVECT_LF) launch_stage_kind<STAGE_ADVECT_LF>(c.stream(), a, tl.cfg, (int)nblk);
  else launch_stage_kind<STAGE_REINIT>(c.stream(), a, tl.cfg, (int)nblk);
  c.launched("k_stage");
  if (fill_after) fill_ghost_box(c, out, out_gb, fbp, bc);
  c.finish();
}

}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" {

void lsm3d_amd_set_stage_tiling(int cfg, int kz) {
  if (cfg < 0 || cfg >= kNumTiles || kz < 1) {
    raise_error("lsm3d_amd_set_stage_tiling: invalid tiling");
    return;
  }
  tl_tiling.cfg = cfg;
  tl_tiling.kz = kz;
}

int lsm3d_amd_stage_tiling_count(void) { return kNumTiles; }

void lsm3d_amd_set_zero_tol(double tol) { tl_zero_tol = tol < 0 ? (double)Tol<real_t>::zero : tol; }

double lsm3d_amd_zero_tol(void) { return tl_zero_tol; }

void lsm3d_amd_stage_tiling_shape(int cfg, int* tx, int* ty) {
  if (cfg < 0 || cfg >= kNumTiles) {
    *tx = *ty = 0;
    return;
  }
  *tx = kTiles[cfg].tx;
  *ty = kTiles[cfg].ty;
}

void lsm3d_amd_advect_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6],
                            const LSMLIB_REAL* phi_stage, const int stage_gb[6], const LSMLIB_REAL* phi_cur,
                            const int cur_gb[6], const LSMLIB_REAL* vel_x, const LSMLIB_REAL* vel_y,
                            const LSMLIB_REAL* vel_z, const int vel_gb[6], const int fb[6], LSMLIB_REAL dx,
                            LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6], int periodic_mask) {
  run_stage("lsm3d_amd_advect_stage", STAGE_ADVECT_FIELD, rk_form, phi_out, out_gb, phi_stage, stage_gb,
            phi_cur, cur_gb, vel_x, vel_y, vel_z, vel_gb, nullptr, nullptr, 0, 0, 0, fb, dx, dy, dz, dt, 0, bc,
            periodic_mask);
}

void lsm3d_amd_advect_const_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6],
                                  const LSMLIB_REAL* phi_stage, const int stage_gb[6],
                                  const LSMLIB_REAL* phi_cur, const int cur_gb[6], LSMLIB_REAL vel_x,
                                  LSMLIB_REAL vel_y, LSMLIB_REAL vel_z, const int fb[6], LSMLIB_REAL dx,
                                  LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6],
                                  int periodic_mask) {
  run_stage("lsm3d_amd_advect_const_stage", STAGE_ADVECT_CONST, rk_form, phi_out, out_gb, phi_stage,
            stage_gb, phi_cur, cur_gb, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, vel_x, vel_y,
            vel_z, fb, dx, dy, dz, dt, 0, bc, periodic_mask);
}

void lsm3d_amd_advect_lf_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6],
                               const LSMLIB_REAL* phi_stage, const int stage_gb[6], const LSMLIB_REAL* phi_cur,
                               const int cur_gb[6], const LSMLIB_REAL* vel_x, const LSMLIB_REAL* vel_y,
                               const LSMLIB_REAL* vel_z, const int vel_gb[6], const int fb[6], LSMLIB_REAL dx,
                               LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6], int periodic_mask,
                               int alpha_mode, const LSMLIB_REAL* alpha) {
  run_stage("lsm3d_amd_advect_lf_stage", STAGE_ADVECT_LF, rk_form, phi_out, out_gb, phi_stage, stage_gb, phi_cur,
            cur_gb, vel_x, vel_y, vel_z, vel_gb, nullptr, nullptr, 0, 0, 0, fb, dx, dy, dz, dt, 0, bc, periodic_mask,
            alpha_mode, alpha);
}

void lsm3d_amd_reinit_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6],
                            const LSMLIB_REAL* phi_stage, const int stage_gb[6], const LSMLIB_REAL* phi_cur,
                            const int cur_gb[6], const LSMLIB_REAL* phi0, const int phi0_gb[6],
                            int use_phi0_for_sgn, const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy,
                            LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6], int periodic_mask) {
  run_stage("lsm3d_amd_reinit_stage", STAGE_REINIT, rk_form, phi_out, out_gb, phi_stage, stage_gb, phi_cur,
            cur_gb, nullptr, nullptr, nullptr, nullptr, phi0, phi0_gb, 0, 0, 0, fb, dx, dy, dz, dt,
            use_phi0_for_sgn == 1, bc, periodic_mask);
}

void lsm3d_amd_fill_ghosts(LSMLIB_REAL* phi, const int gb[6], const int fb[6], const int bc[6]) {
  const char* fn = "lsm3d_amd_fill_ghosts";
  Call c(fn);
  const Box g = Box{gb[0], gb[1], gb[2], gb[3], gb[4], gb[5]};
  const Box f = Box{fb[0], fb[1], fb[2], fb[3], fb[4], fb[5]};
  require_inside(fn, "phi", g, f, 0, 0, 0);
  const int n[3] = {f.nx(), f.ny(), f.nz()};
  const int gw[3][2] = {{f.ilo - g.ilo, g.ihi - f.ihi}, {f.jlo - g.jlo, g.jhi - f.jhi}, {f.klo - g.klo, g.khi - f.khi}};
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      const int code = bc[2 * d + s];
      if (code != 0 && code != 1 && code != 2 && code != 3 && code != 5)
        raise_error(std::string(fn) + ": unknown boundary condition code");
      if ((code == 2 || code == 3) && n[d] < 2) raise_error(std::string(fn) + ": extrapolation needs 2 cells");
      if (code == 5 && gw[d][s] > n[d]) raise_error(std::string(fn) + ": periodic ghost wider than the box");
    }
  device_only(fn, phi, "phi");
  if (c.ok()) fill_ghost_box(c, phi, gb, fb, bc);
  c.finish();
}

}  // extern "C"
