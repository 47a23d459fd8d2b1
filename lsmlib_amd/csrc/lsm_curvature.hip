// Fused stages of LSMLIB's "medium accuracy" model path on device-resident
// fields: second-order HJ-ENO upwinding, TVD-RK2 in time, the constant
// normal-velocity and mean-curvature terms, and the 'phi <= mask' constraint.
//
// One launch advances one stage over the whole fill box:
//   curvature:  u_out = RK(u_cur, u_s, -vn |grad u_s|_Godunov + b kappa |grad u_s|)
//               the sequence LSM3D_HJ_ENO2 -> LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS
//               -> 4 x LSM3D_CENTRAL_GRAD_ORDER2 -> LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS
//               -> LSM3D_TVD_RK2_STAGE* of the curvature example
//               (examples/serial/curvature_example/curvature_model3d.c:134-417)
//   reinit:     u_out = RK(u_cur, u_s, S(u)(1 - |grad u_s|*)) with HJ-ENO2
//               (curvature_model3d.c:484-625)
// followed, when a mask is given, by IMPOSE_MASK (src/serial/lsm_macros.h:84-91)
// on the fill box.  The reference fills the ghost cells by signed linear
// extrapolation from the UNMASKED stage result and masks afterwards
// (curvature_model3d.c:293-296, 410-414); the stage therefore also keeps the
// unmasked values of the two outermost fill-box layers in `raw`, and the
// ghost kernel below extrapolates from those and masks the ghost cells.
//
// The stage kernel marches each (TX x TY) column tile through its z chunk
// with a six-plane LDS ring: a plane with a two-cell x/y halo enters the ring
// once per step, every stencil value (ENO2 +-2 along each axis, the central
// second and mixed derivatives) is read from LDS, and one barrier per plane
// suffices because the plane written at step k (k+2) never aliases the five
// planes read at steps k-1 and k.
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_march.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

namespace lsm {

enum MediumKind { MED_CURVATURE = 0, MED_REINIT = 1, MED_DT = 2 };

template <typename T>
struct MediumArgs {
  Field<T> out;         // stage result on the fill box (masked when mask.p)
  Field<T> raw;         // unmasked result on the two outermost fill-box layers
  Field<const T> ps;    // stage input, ghost cells valid two cells out
  Field<const T> pc;    // u_cur of TVD-RK2 stage 2
  Field<const T> mask;  // p == nullptr: no mask
  Field<const T> phi0;  // reinitialization sign field (use_phi0)
  Box fb;
  TileGrid tg;
  T ix, iy, iz;  // 1/dx: ENO2 derivatives
  T hx, hy, hz;  // 0.5d0/dx rounded to T: central differences (lsm_spatial_derivatives3d.f.in:3058-3060)
  T vn;          // constant normal speed (no term when |vn| < zero_tol)
  T b;           // curvature coefficient (no term when 0)
  T dt, dx_sq, zero_tol;  // MED_DT: vn = |vel_n|, dx_sq = max_dx^2
  int form, use_phi0;
  double* partials;       // MED_DT: one max per workgroup
};

template <typename T, int TX, int TY, int KIND>
__global__ void __launch_bounds__(TX* TY) k_medium_stage(MediumArgs<T> a) {
#pragma clang fp contract(off)  // every product and sum rounded as the reference's
  constexpr int W = TX + 4, H = TY + 4, NT = TX * TY, PER = (W * H + NT - 1) / NT;
  __shared__ T ring[6][H][W + 1];
  const int blk = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  int bx, by, bz;
  tile_coords(a.tg, blk, bx, by, bz);
  const Box& fb = a.fb;
  const int t = (int)threadIdx.x, tx = t % TX, ty = t / TX;
  const int i0 = fb.ilo + bx * TX, j0 = fb.jlo + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = fb.klo + bz * a.tg.kz;
  const int k1 = min(k0 + a.tg.kz - 1, fb.khi);
  const bool active = (i <= fb.ihi) && (j <= fb.jhi);

  // This thread's cells of a plane's (W x H) window: LDS position and the
  // element offset inside a plane of phi_stage (the same for every plane).
  // Lanes past the fill box of a partial tile load clamped (unused) cells.
  constexpr int SLOT = H * (W + 1);  // elements per ring slot
  T* lds = &ring[0][0][0];
  int lpos[PER], goff[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int q = min(t + r * NT, W * H - 1);
    const int qy = q / W, qx = q - qy * W;
    lpos[r] = qy * (W + 1) + qx;
    const int gi = min(i0 - 2 + qx, fb.ihi + 2), gj = min(j0 - 2 + qy, fb.jhi + 2);
    goff[r] = (gi - a.ps.ilo) + (int)a.ps.sy * (gj - a.ps.jlo);
  }
  // In-plane offsets of the cell in the other fields (plane bases advance with k).
  const int cpos = (ty + 2) * (W + 1) + (tx + 2);
  const int o_out = (i - a.out.ilo) + (int)a.out.sy * (j - a.out.jlo);
  const int o_raw = (i - a.raw.ilo) + (int)a.raw.sy * (j - a.raw.jlo);
  // AHEAD (the reinitialization stage): the mask, phi_cur and phi0 values of
  // a plane are loaded one step ahead and unconditionally -- lanes past the
  // fill box read a clamped cell, an absent field reads the stage field's
  // cell (unused) -- because a load under a runtime branch, or one consumed
  // in the step that issued it, makes the compiler wait for the next plane's
  // prefetch as well (-15 % per reinitialization stage at 512 x 512 x 341,
  // FP32).  The curvature stage, whose arithmetic is longer, measured 2 %
  // slower that way and keeps loading plane k's values after the prefetch.
  constexpr bool AHEAD = KIND == MED_REINIT;
  const int o_mask = (i - a.mask.ilo) + (int)a.mask.sy * (j - a.mask.jlo);
  const int o_pc = (i - a.pc.ilo) + (int)a.pc.sy * (j - a.pc.jlo);
  T mk = T(0), uk = T(0), pk = T(0);  // AHEAD: the next step's plane values
  int om = 0, ou = 0, op = 0;          // AHEAD: in-plane offsets of the three loads
  if constexpr (AHEAD) {
    const int ic = min(i, fb.ihi), jc = min(j, fb.jhi);
    const int o_ps = (ic - a.ps.ilo) + (int)a.ps.sy * (jc - a.ps.jlo);
    om = a.mask.p ? (ic - a.mask.ilo) + (int)a.mask.sy * (jc - a.mask.jlo) : o_ps;
    ou = a.form == RK2_STAGE2 ? (ic - a.pc.ilo) + (int)a.pc.sy * (jc - a.pc.jlo) : o_ps;
    op = a.use_phi0 ? (ic - a.phi0.ilo) + (int)a.phi0.sy * (jc - a.phi0.jlo) : o_ps;
  }
  auto plane_of = [&](const Field<const T>& f, bool has, int kk) -> const T* {
    return has ? f.p + f.sz * (int64_t)(kk - f.klo) : a.ps.p + a.ps.sz * (int64_t)(kk - a.ps.klo);
  };
  auto side = [&](int kk) {  // AHEAD: loads of plane kk
    mk = __builtin_nontemporal_load(plane_of(a.mask, a.mask.p != nullptr, kk) + om);
    uk = __builtin_nontemporal_load(plane_of(a.pc, a.form == RK2_STAGE2, kk) + ou);
    pk = __builtin_nontemporal_load(plane_of(a.phi0, a.use_phi0 != 0, kk) + op);
  };
  T nxt[PER];
  auto fetch = [&](int kk) {  // plane kk into registers
    const T* base = a.ps.p + a.ps.sz * (int64_t)(kk - a.ps.klo);
#pragma unroll
    for (int r = 0; r < PER; r++) nxt[r] = base[goff[r]];
  };
  auto commit = [&](int slot) {  // registers into ring slot `slot`
    T* dst = lds + slot * SLOT;
#pragma unroll
    for (int r = 0; r < PER; r++)
      if (r < PER - 1 || t + r * NT < W * H) dst[lpos[r]] = nxt[r];
  };
  auto wrap6 = [](int v) { return v >= 6 ? v - 6 : (v < 0 ? v + 6 : v); };

  double acc = -1.0e308;  // MED_DT: running max of the CFL speed
  for (int kk = k0 - 2; kk <= k0 + 1; kk++) {  // plane kk -> slot kk - k0 + 2
    fetch(kk);
    commit(kk - k0 + 2);
  }
  if constexpr (AHEAD) side(k0);
  fetch(k0 + 2);
  int sc = 2;  // ring slot of plane k (uniform)
  for (int k = k0; k <= k1; k++, sc = wrap6(sc + 1)) {
    sc = __builtin_amdgcn_readfirstlane(sc);
    // Plane k+2 enters the ring (slot k-4 mod 6, read last at step k-2) and
    // plane k+3 is fetched; its loads stay in flight over this step's
    // arithmetic.  The one barrier orders the ring writes before the reads.
    commit(wrap6(sc + 2));
    T m = T(0), uc = T(0), p0 = T(0);
    if constexpr (AHEAD) {
      // plane k's values (landed: loaded before plane k+2); plane k+1's are
      // fetched before plane k+3 (at the last steps planes k1 and k1+2 are
      // re-read, unused)
      m = mk;
      uc = uk;
      p0 = pk;
      side(min(k + 1, k1));
      fetch(min(k + 3, k1 + 2));
    } else {
      if (k + 3 <= k1 + 2) fetch(k + 3);
      if (active && KIND != MED_DT) {
        if (a.mask.p) m = __builtin_nontemporal_load(a.mask.p + a.mask.sz * (int64_t)(k - a.mask.klo) + o_mask);
        if (a.form == RK2_STAGE2) uc = __builtin_nontemporal_load(a.pc.p + a.pc.sz * (int64_t)(k - a.pc.klo) + o_pc);
      }
    }
    __syncthreads();
    if (active) {
      // slot bases of planes k-2 .. k+2
      const int sb[5] = {wrap6(sc - 2) * SLOT + cpos, wrap6(sc - 1) * SLOT + cpos, sc * SLOT + cpos,
                         wrap6(sc + 1) * SLOT + cpos, wrap6(sc + 2) * SLOT + cpos};
      // P(dz, dy, dx): u_s at (i+dx, j+dy, k+dz)
      auto P = [&](int dz, int dy, int dx) -> T { return lds[sb[dz + 2] + dy * (W + 1) + dx]; };
      T p[7];
#pragma unroll
      for (int q = -2; q <= 2; q++) p[3 + q] = P(0, 0, q);
      const T xp = hj_plus<T, ENO2>(p, a.ix), xm = hj_minus<T, ENO2>(p, a.ix);
#pragma unroll
      for (int q = -2; q <= 2; q++) p[3 + q] = P(0, q, 0);
      const T yp = hj_plus<T, ENO2>(p, a.iy), ym = hj_minus<T, ENO2>(p, a.iy);
#pragma unroll
      for (int q = -2; q <= 2; q++) p[3 + q] = P(q, 0, 0);
      const T zp = hj_plus<T, ENO2>(p, a.iz), zm = hj_minus<T, ENO2>(p, a.iz);
      const T us = P(0, 0, 0);
      if (KIND == MED_DT) {  // lsm_utilities3d.f.in:530-560
        const T h = h_over_dx<T>(a.vn, xp, yp, zp, xm, ym, zm, a.dx_sq, a.ix, a.iy, a.iz);
        acc = acc > (double)h ? acc : (double)h;
      } else {
        T r = T(0);
        if (KIND == MED_CURVATURE) {
          if (fabs(a.vn) >= a.zero_tol)  // lsm_level_set_evolution3d.f.in:300-330
            r = r - a.vn * sqrt(godunov_norm_sq<T>(a.vn, xp, yp, zp, xm, ym, zm));
          if (a.b != T(0)) {
            // First derivatives at the cell and its neighbours (one
            // LSM3D_CENTRAL_GRAD_ORDER2 of u), then the second derivatives as
            // central differences of those: xx, xy, xz from d/dx of phi_x,
            // phi_y, phi_z; yy, yz from d/dy of phi_y, phi_z; zz from d/dz of
            // phi_z (the last writer of each array in curvature_model3d.c:204-243).
            const T gx = (P(0, 0, 1) - P(0, 0, -1)) * a.hx;
            const T gy = (P(0, 1, 0) - P(0, -1, 0)) * a.hy;
            const T gz = (P(1, 0, 0) - P(-1, 0, 0)) * a.hz;
            const T gx_ip = (P(0, 0, 2) - us) * a.hx, gx_im = (us - P(0, 0, -2)) * a.hx;
            const T gy_jp = (P(0, 2, 0) - us) * a.hy, gy_jm = (us - P(0, -2, 0)) * a.hy;
            const T gz_kp = (P(2, 0, 0) - us) * a.hz, gz_km = (us - P(-2, 0, 0)) * a.hz;
            const T gy_ip = (P(0, 1, 1) - P(0, -1, 1)) * a.hy, gy_im = (P(0, 1, -1) - P(0, -1, -1)) * a.hy;
            const T gz_ip = (P(1, 0, 1) - P(-1, 0, 1)) * a.hz, gz_im = (P(1, 0, -1) - P(-1, 0, -1)) * a.hz;
            const T gz_jp = (P(1, 1, 0) - P(-1, 1, 0)) * a.hz, gz_jm = (P(1, -1, 0) - P(-1, -1, 0)) * a.hz;
            const T xx = (gx_ip - gx_im) * a.hx, yy = (gy_jp - gy_jm) * a.hy, zz = (gz_kp - gz_km) * a.hz;
            const T xy = (gy_ip - gy_im) * a.hx, xz = (gz_ip - gz_im) * a.hx, yz = (gz_jp - gz_jm) * a.hy;
            r = r + a.b * curvature_term<T>(gx, gy, gz, xx, xy, xz, yy, yz, zz, a.zero_tol);
          }
        } else {  // lsm_reinitialization3d.f.in:160-200
          const T cur = a.use_phi0 ? p0 : us;  // this branch is the AHEAD kind
          if (fabs(cur) >= a.zero_tol) {
            const T n2 = godunov_star_sq<T>(cur, xp, yp, zp, xm, ym, zm);
            const T sgn = a.use_phi0 ? cur / sqrt(cur * cur + a.dx_sq) : cur / sqrt(cur * cur + n2 * a.dx_sq);
            r = sgn * (T(1) - sqrt(n2));
          }
        }
        const T y = rk_combine<T>(a.form, (a.form == RK2_STAGE2) ? uc : us, us, r, a.dt);
        T o = y;
        if (a.mask.p) {
          o = (m > y) ? m : y;
          const bool shell = (i - fb.ilo < 2) || (fb.ihi - i < 2) || (j - fb.jlo < 2) || (fb.jhi - j < 2) ||
                             (k - fb.klo < 2) || (fb.khi - k < 2);
          if (shell) a.raw.p[a.raw.sz * (int64_t)(k - a.raw.klo) + o_raw] = y;
        }
        __builtin_nontemporal_store(o, a.out.p + a.out.sz * (int64_t)(k - a.out.klo) + o_out);
      }
    }
  }
  if (KIND == MED_DT) {  // block max -> one partial per workgroup
    __shared__ double s_w[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_xor(acc, o, 64);
      acc = acc > w ? acc : w;
    }
    if ((t & 63) == 0) s_w[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      double r = s_w[0];
      for (int w = 1; w < NT / 64; w++) r = r > s_w[w] ? r : s_w[w];
      a.partials[blockIdx.x] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Ghost cells of the serial package's boundary fill for all six faces
// (signedLinearExtrapolationBC with bdry_location_idx 9,
// src/serial/lsm_boundary_conditions.c:306-575: faces x_lo, x_hi, y_lo, y_hi,
// z_lo, z_hi in that order, each over the whole ghost box in the other two
// directions).  Every ghost cell is written once with the value that sequence
// leaves in it: a z-face cell extrapolates the two values the x and y faces
// left next to it, a y-face cell the two the x faces left, an x-face cell two
// fill-box cells.  The fill-box values come from `src` (the unmasked stage
// result, or phi itself); with a mask the cell takes max(mask, value).
// Codes per face: 1 copy, 2 linear, 3 signed linear.
struct FaceCodes {
  int code[6];
};

template <typename T>
__device__ __forceinline__ T ghost_x(const Field<const T>& s, const Box& fb, const FaceCodes& c, int i, int j,
                                     int k) {
  if (i < fb.ilo) return extrapolate<T>(c.code[0], s.at(fb.ilo, j, k), s.at(fb.ilo + 1, j, k), fb.ilo - i);
  if (i > fb.ihi) return extrapolate<T>(c.code[1], s.at(fb.ihi, j, k), s.at(fb.ihi - 1, j, k), i - fb.ihi);
  return s.at(i, j, k);
}

template <typename T>
__device__ __forceinline__ T ghost_xy(const Field<const T>& s, const Box& fb, const FaceCodes& c, int i, int j,
                                      int k) {
  if (j < fb.jlo)
    return extrapolate<T>(c.code[2], ghost_x<T>(s, fb, c, i, fb.jlo, k), ghost_x<T>(s, fb, c, i, fb.jlo + 1, k),
                          fb.jlo - j);
  if (j > fb.jhi)
    return extrapolate<T>(c.code[3], ghost_x<T>(s, fb, c, i, fb.jhi, k), ghost_x<T>(s, fb, c, i, fb.jhi - 1, k),
                          j - fb.jhi);
  return ghost_x<T>(s, fb, c, i, j, k);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_ghosts_masked(Field<T> dst, Field<const T> src, Field<const T> mask, Box gb, Box fb, FaceCodes c,
                    int64_t n_z, int64_t n_y, int64_t n_all) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_all) return;
  const int gnx = gb.nx(), gny = gb.ny();
  const int xl = fb.ilo - gb.ilo, yl = fb.jlo - gb.jlo, zl = fb.klo - gb.klo;
  const int xw = gnx - fb.nx(), yw = gny - fb.ny();
  int i, j, k;
  if (q < n_z) {  // below and above the fill box in z: whole (x, y) planes
    const int64_t plane = (int64_t)gnx * gny;
    const int layer = (int)(q / plane);
    const int64_t rem = q - (int64_t)layer * plane;
    k = layer < zl ? gb.klo + layer : fb.khi + 1 + (layer - zl);
    j = gb.jlo + (int)(rem / gnx);
    i = gb.ilo + (int)(rem - (rem / gnx) * gnx);
  } else if (q < n_z + n_y) {  // y ghost rows of the fill-box planes
    const int64_t r = q - n_z, per = (int64_t)gnx * yw;
    const int kk = (int)(r / per);
    const int64_t rem = r - (int64_t)kk * per;
    const int layer = (int)(rem / gnx);
    k = fb.klo + kk;
    j = layer < yl ? gb.jlo + layer : fb.jhi + 1 + (layer - yl);
    i = gb.ilo + (int)(rem - (int64_t)layer * gnx);
  } else {  // x ghost cells of the fill-box rows
    const int64_t r = q - n_z - n_y;
    const int64_t line = r / xw;
    const int layer = (int)(r - line * xw);
    k = fb.klo + (int)(line / fb.ny());
    j = fb.jlo + (int)(line - (line / fb.ny()) * fb.ny());
    i = layer < xl ? gb.ilo + layer : fb.ihi + 1 + (layer - xl);
  }
  T v;
  if (k < fb.klo)
    v = extrapolate<T>(c.code[4], ghost_xy<T>(src, fb, c, i, j, fb.klo), ghost_xy<T>(src, fb, c, i, j, fb.klo + 1),
                       fb.klo - k);
  else if (k > fb.khi)
    v = extrapolate<T>(c.code[5], ghost_xy<T>(src, fb, c, i, j, fb.khi), ghost_xy<T>(src, fb, c, i, j, fb.khi - 1),
                       k - fb.khi);
  else
    v = ghost_xy<T>(src, fb, c, i, j, k);
  if (mask.p) {
    const T m = mask.at(i, j, k);
    v = (m > v) ? m : v;
  }
  dst.at(i, j, k) = v;
}

// ---------------------------------------------------------------------------
// Stable dt of the constant normal velocity (MED_DT above: the HJ-ENO2
// derivatives of phi formed in the stage kernel's LDS ring, then
// LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT's max over the fill box,
// lsm_utilities3d.f.in:455-565): per-workgroup maxima, then one workgroup.
__global__ void __launch_bounds__(256) k_max_partials(const double* partials, int64_t n, double* out) {
  __shared__ double s_w[4];
  double acc = -1.0e308;
  for (int64_t q = threadIdx.x; q < n; q += blockDim.x) acc = acc > partials[q] ? acc : partials[q];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(acc, o, 64);
    acc = acc > w ? acc : w;
  }
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_w[0];
    for (int w = 1; w < (int)(blockDim.x + 63) / 64; w++) r = r > s_w[w] ? r : s_w[w];
    out[0] = r;
  }
}

// IMPOSE_MASK over a flat array (src/serial/lsm_macros.h:84-91)
template <typename T>
__global__ void __launch_bounds__(256) k_impose_mask(T* out, const T* mask, const T* phi, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const T m = mask[q], v = phi[q];
    out[q] = (m > v) ? m : v;
  }
}

}  // namespace lsm

// ===========================================================================
// Host side
// ===========================================================================
namespace lsm {
Tiling current_tiling();
double current_zero_tol();
TileShape tile_shape(int cfg);
bool valid_form(int form);
bool device_only(const char* fn, const void* p, const char* what);
Field<const real_t> cfield(const real_t* p, const int* gb);
}  // namespace lsm

namespace {
using namespace lsm;
typedef real_t R;

Box box6(const int* b) { return Box{b[0], b[1], b[2], b[3], b[4], b[5]}; }

template <int TX, int TY>
void launch_medium(hipStream_t s, int kind, const MediumArgs<R>& a, int nblk) {
  if (kind == MED_CURVATURE)
    hipLaunchKernelGGL((k_medium_stage<R, TX, TY, MED_CURVATURE>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
  else if (kind == MED_REINIT)
    hipLaunchKernelGGL((k_medium_stage<R, TX, TY, MED_REINIT>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
  else
    hipLaunchKernelGGL((k_medium_stage<R, TX, TY, MED_DT>), dim3((unsigned)nblk), dim3(TX * TY), 0, s, a);
}

void launch_medium_cfg(hipStream_t s, int cfg, int kind, const MediumArgs<R>& a, int nblk) {
  switch (cfg) {
    case 0: launch_medium<64, 4>(s, kind, a, nblk); break;
    case 1: launch_medium<64, 8>(s, kind, a, nblk); break;
    case 2: launch_medium<32, 8>(s, kind, a, nblk); break;
    case 3: launch_medium<128, 2>(s, kind, a, nblk); break;
    case 4: launch_medium<128, 4>(s, kind, a, nblk); break;
    case 5: launch_medium<32, 16>(s, kind, a, nblk); break;
    case 6: launch_medium<16, 16>(s, kind, a, nblk); break;
    case 7: launch_medium<16, 8>(s, kind, a, nblk); break;
    default: launch_medium<16, 4>(s, kind, a, nblk); break;
  }
}

// Common fields of the three kinds; false (error raised) when the launch
// would not fit.
bool medium_args(const char* fn, MediumArgs<R>& a, const Box& fb, const R* ps, const int* ps_gb, R dx, R dy,
                 R dz, long long& nblk, int& cfg) {
  const Tiling tl = current_tiling();
  const TileShape ts = tile_shape(tl.cfg);
  const Box g = box6(ps_gb);
  if ((long long)g.nx() * g.ny() > 0x7fffffffLL) {
    raise_error(std::string(fn) + ": phi_stage planes above 2^31 elements");
    return false;
  }
  a.ps = cfield(ps, ps_gb);
  a.fb = fb;
  a.tg = make_tile_grid(fb, ts.tx, ts.ty, tl.kz);
  a.ix = R(1) / dx;
  a.iy = R(1) / dy;
  a.iz = R(1) / dz;
  a.hx = R(0.5 / (double)dx);
  a.hy = R(0.5 / (double)dy);
  a.hz = R(0.5 / (double)dz);
  R mdx = dx;
  if (dy > mdx) mdx = dy;
  if (dz > mdx) mdx = dz;
  a.dx_sq = mdx * mdx;  // lsm_reinitialization3d.f.in:140-141, lsm_utilities3d.f.in:518-519
  a.zero_tol = (R)current_zero_tol();
  a.partials = nullptr;
  nblk = (long long)a.tg.ntx * a.tg.nty * a.tg.nkz;
  cfg = tl.cfg;
  if (nblk > 0x7fffffffLL) {
    raise_error(std::string(fn) + ": grid too large for one launch");
    return false;
  }
  return true;
}

void medium_stage(const char* fn, int kind, int form, R* out, const int* out_gb, R* raw, const int* raw_gb,
                  const R* ps, const int* ps_gb, const R* pc, const int* pc_gb, const R* mask, const int* mask_gb,
                  const R* phi0, const int* phi0_gb, int use_phi0, const int* fbp, R dx, R dy, R dz, R vn, R b,
                  R dt) {
  Call c(fn);
  const Box fb = box6(fbp);
  if (form != RK_EULER && form != RK2_STAGE2) raise_error(std::string(fn) + ": rk_form must be 1 or 22");
  require_inside(fn, "phi_out", box6(out_gb), fb, 0, 0, 0);
  require_inside(fn, "phi_stage", box6(ps_gb), fb, 2, 2, 2);
  if (form == RK2_STAGE2) {
    if (!pc) raise_error(std::string(fn) + ": TVD-RK2 stage 2 needs phi_cur");
    else require_inside(fn, "phi_cur", box6(pc_gb), fb, 0, 0, 0);
  }
  if (mask) {
    require_inside(fn, "mask", box6(mask_gb), fb, 0, 0, 0);
    if (!raw) raise_error(std::string(fn) + ": a mask needs the unmasked-result array phi_raw");
    else require_inside(fn, "phi_raw", box6(raw_gb), fb, 0, 0, 0);
  }
  if (kind == MED_REINIT && use_phi0) require_inside(fn, "phi0", box6(phi0_gb), fb, 0, 0, 0);
  // the kernel keeps in-plane offsets in 32 bits
  const int* gbs[5] = {out_gb, raw ? raw_gb : nullptr, pc ? pc_gb : nullptr, mask ? mask_gb : nullptr,
                       (kind == MED_REINIT && use_phi0) ? phi0_gb : nullptr};
  for (const int* g : gbs)
    if (g && (long long)box6(g).nx() * box6(g).ny() > 0x7fffffffLL)
      raise_error(std::string(fn) + ": array planes above 2^31 elements");
  if (!(device_only(fn, out, "phi_out") && device_only(fn, raw, "phi_raw") && device_only(fn, ps, "phi_stage") &&
        device_only(fn, pc, "phi_cur") && device_only(fn, mask, "mask") && device_only(fn, phi0, "phi0")))
    return;
  if (!c.ok() || fb.cells() == 0) {
    c.finish();
    return;
  }
  MediumArgs<R> a;
  long long nblk = 0;
  int cfg = 0;
  if (!medium_args(fn, a, fb, ps, ps_gb, dx, dy, dz, nblk, cfg)) {
    c.finish();
    return;
  }
  a.out = make_field<R>(out, box6(out_gb));
  a.raw = mask ? make_field<R>(raw, box6(raw_gb)) : make_field<R>(nullptr, fb);
  a.pc = cfield(form == RK2_STAGE2 ? pc : nullptr, pc_gb);
  a.mask = cfield(mask, mask_gb);
  a.phi0 = cfield((kind == MED_REINIT && use_phi0) ? phi0 : nullptr, phi0_gb);
  a.vn = vn;
  a.b = b;
  a.dt = dt;
  a.form = form;
  a.use_phi0 = (kind == MED_REINIT && use_phi0) ? 1 : 0;
  launch_medium_cfg(c.stream(), cfg, kind, a, (int)nblk);
  c.launched("k_medium_stage");
  c.finish();
}

}  // namespace

extern "C" {

void lsm3d_amd_curvature_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6], LSMLIB_REAL* phi_raw,
                               const int raw_gb[6], const LSMLIB_REAL* phi_stage, const int stage_gb[6],
                               const LSMLIB_REAL* phi_cur, const int cur_gb[6], const LSMLIB_REAL* mask,
                               const int mask_gb[6], const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy,
                               LSMLIB_REAL dz, LSMLIB_REAL vel_n, LSMLIB_REAL b, LSMLIB_REAL dt) {
  medium_stage("lsm3d_amd_curvature_stage", MED_CURVATURE, rk_form, phi_out, out_gb, phi_raw, raw_gb, phi_stage,
               stage_gb, phi_cur, cur_gb, mask, mask_gb, nullptr, nullptr, 0, fb, dx, dy, dz, vel_n, b, dt);
}

void lsm3d_amd_reinit_eno2_stage(int rk_form, LSMLIB_REAL* phi_out, const int out_gb[6], LSMLIB_REAL* phi_raw,
                                 const int raw_gb[6], const LSMLIB_REAL* phi_stage, const int stage_gb[6],
                                 const LSMLIB_REAL* phi_cur, const int cur_gb[6], const LSMLIB_REAL* phi0,
                                 const int phi0_gb[6], int use_phi0_for_sgn, const LSMLIB_REAL* mask,
                                 const int mask_gb[6], const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy,
                                 LSMLIB_REAL dz, LSMLIB_REAL dt) {
  medium_stage("lsm3d_amd_reinit_eno2_stage", MED_REINIT, rk_form, phi_out, out_gb, phi_raw, raw_gb, phi_stage,
               stage_gb, phi_cur, cur_gb, mask, mask_gb, phi0, phi0_gb, use_phi0_for_sgn == 1, fb, dx, dy, dz,
               R(0), R(0), dt);
}

void lsm3d_amd_extrapolate_ghosts_masked(LSMLIB_REAL* phi, const int gb[6], const LSMLIB_REAL* src,
                                         const int src_gb[6], const LSMLIB_REAL* mask, const int mask_gb[6],
                                         const int fb[6], const int bc[6]) {
  const char* fn = "lsm3d_amd_extrapolate_ghosts_masked";
  Call c(fn);
  const Box g = box6(gb), f = box6(fb);
  require_inside(fn, "phi", g, f, 0, 0, 0);
  if (f.nx() < 2 || f.ny() < 2 || f.nz() < 2) raise_error(std::string(fn) + ": the fill box needs 2 cells per axis");
  FaceCodes fc;
  for (int q = 0; q < 6; q++) {
    if (bc[q] < 1 || bc[q] > 3) raise_error(std::string(fn) + ": face codes must be 1 (copy), 2 or 3 (extrapolation)");
    fc.code[q] = bc[q];
  }
  const R* s = src ? src : phi;
  const int* sgb = src ? src_gb : gb;
  require_inside(fn, "src", box6(sgb), f, 0, 0, 0);
  if (mask) require_inside(fn, "mask", box6(mask_gb), g, 0, 0, 0);
  if (!(device_only(fn, phi, "phi") && device_only(fn, src, "src") && device_only(fn, mask, "mask"))) return;
  const int64_t n_z = (int64_t)g.nx() * g.ny() * (g.nz() - f.nz());
  const int64_t n_y = (int64_t)g.nx() * (g.ny() - f.ny()) * f.nz();
  const int64_t n_x = (int64_t)(g.nx() - f.nx()) * f.ny() * f.nz();
  const int64_t n_all = n_z + n_y + n_x;
  if (c.ok() && n_all > 0) {
    const int64_t nb = (n_all + 255) / 256;
    hipLaunchKernelGGL(k_ghosts_masked<R>, dim3((unsigned)nb), dim3(256), 0, c.stream(), make_field<R>(phi, g),
                       make_field<const R>(s, box6(sgb)), cfield(mask, mask_gb), g, f, fc, n_z, n_y, n_all);
    c.launched("k_ghosts_masked");
  }
  c.finish();
}

void lsm3d_amd_const_normal_vel_dt_eno2(LSMLIB_REAL* dt, const LSMLIB_REAL* phi, const int phi_gb[6],
                                        const int fb[6], LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz,
                                        LSMLIB_REAL vel_n, LSMLIB_REAL cfl) {
  const char* fn = "lsm3d_amd_const_normal_vel_dt_eno2";
  Call c(fn);
  const Box f = box6(fb);
  require_inside(fn, "phi", box6(phi_gb), f, 2, 2, 2);
  if (!device_only(fn, phi, "phi")) return;
  double m = -1.0e308;
  MediumArgs<R> a{};
  long long nblk = 0;
  int cfg = 0;
  if (c.ok() && f.cells() > 0 && medium_args(fn, a, f, phi, phi_gb, dx, dy, dz, nblk, cfg)) {
    double* scratch = reinterpret_cast<double*>(device_scratch((size_t)(nblk + 1) * sizeof(double)));
    if (c.ok()) {
      a.vn = R(std::fabs(vel_n));
      a.partials = scratch + 1;
      launch_medium_cfg(c.stream(), cfg, MED_DT, a, (int)nblk);
      c.launched("k_medium_stage(dt)");
      hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, c.stream(), scratch + 1, (int64_t)nblk, scratch);
      c.launched("k_max_partials");
      check_hip(hipMemcpyAsync(&m, scratch, sizeof(double), hipMemcpyDeviceToHost, c.stream()),
                "hipMemcpyAsync(dt)");
      check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
    }
  }
  c.finish();
  R mx = R(-1.0);  // lsm_utilities3d.f.in:523 (max starts at -1)
  if (m > (double)mx) mx = R(m);
  if (!lsm::error_pending()) *dt = cfl / (mx + Tol<R>::tiny);
}

void lsm3d_amd_impose_mask(LSMLIB_REAL* phi_masked, const LSMLIB_REAL* mask, const LSMLIB_REAL* phi,
                           long long num_gridpts) {
  const char* fn = "lsm3d_amd_impose_mask";
  Call c(fn);
  if (num_gridpts < 0) raise_error(std::string(fn) + ": negative size");
  if (!(device_only(fn, phi_masked, "phi_masked") && device_only(fn, mask, "mask") && device_only(fn, phi, "phi")))
    return;
  if (c.ok() && num_gridpts > 0) {
    int64_t nb = (num_gridpts + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_impose_mask<R>, dim3((unsigned)nb), dim3(256), 0, c.stream(), phi_masked, mask, phi,
                       (int64_t)num_gridpts);
    c.launched("k_impose_mask");
  }
  c.finish();
}

}  // extern "C"
