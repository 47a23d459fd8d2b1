// Reverse mode of the fused WENO5 advection stage, and the smoothed
// Heaviside mismatch used as a loss on the zero level set.
//
// Forward stage (lsm_fused.hip, the reference path LSM3D_UPWIND_HJ_WENO5 ->
// LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS -> LSM3D_TVD_RK3_STAGE*):
//   out = a_c u_c + a_s (u_s + dt r(u_s, V)),   r = -sum_a V_a D_a(u_s)
// with D_a the upwind HJ-WENO5 derivative along axis a (backward stencil for
// V_a > 0, forward for V_a < 0, zero for |V_a| < zero_tol) and
//   RK1: a_s = 1, a_c = 0         TVD-RK2 stage 2: a_s = a_c = 1/2
//   TVD-RK3 stage 2: a_s = 1/4, a_c = 3/4      stage 3: a_s = 2/3, a_c = 1/3
// (lsm_tvd_runge_kutta3d.f:78, 240-241, 403-404, 486-488).
//
// Given the adjoint ubar of `out`, the reverse stage produces
//   adj_s = a_s (ubar + dt J^T ubar),   J = d r / d u_s
//   adj_c = a_c ubar
//   vbar_a += -a_s dt ubar D_a(u_s)              (d r / d V_a, per cell)
// The selection between the two one-sided stencils and the zero band are
// taken as fixed (piecewise derivative, exact away from V_a = 0 and from
// ties in the max of the WENO epsilon).
//
// J^T is formed in two passes over the fill box, which is treated as
// periodic on every axis:
//   pass 1 (k_adj_faces): per cell m and axis a, the gradient of D_a(m)
//     with respect to its five undivided differences, scaled by
//     -a_s dt ubar(m) V_a(m), is written to six face slots: slot q holds
//     the weight of face m + q - 2 (difference u(m+q-2) - u(m+q-3)).
//   pass 2 (k_adj_gather): ubar_s(i) = sum_a sum_q slot_q(i+2-q) -
//     slot_q(i+3-q) - the transpose of the differencing - plus a_s ubar(i).
// Both passes read u_s, V and the slots with periodic wrap inside the fill
// box, so stored ghost cells are never read.
#include <cstdio>
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

namespace lsm {

// Value F and gradient g[0..4] of the WENO5 combination of weno5_fast (the
// same function as weno5_ref on undivided differences D1..D5; the derivative
// is F/dx).  With a..d the second differences, s_k the scaled smoothness
// indicators plus E = 12e-6 max|D|^2 + tiny_dx2, alpha_k = d_k / s_k^2 and
// w_k = alpha_k / sum alpha:
//   F = p2 + w1 q1/3 + w3 q2/6,   q1 = a - 2b + c,  q2 = b - 2c + d
//   dF/ds_l = -2 w_l (c_l - Phi) / s_l,  c = (q1/3, 0, q2/6),  Phi = F - p2
template <typename T>
__device__ __forceinline__ void weno5_grad(T D1, T D2, T D3, T D4, T D5, T tiny_dx2, T& F, T* g) {
  const T a = D1 - D2, b = D2 - D3, c = D3 - D4, d = D4 - D5;
  const T A1 = fabs(D1), A2 = fabs(D2), A3 = fabs(D3), A4 = fabs(D4), A5 = fabs(D5);
  T m = A1;
  int jm = 0;
  if (A2 > m) { m = A2; jm = 1; }
  if (A3 > m) { m = A3; jm = 2; }
  if (A4 > m) { m = A4; jm = 3; }
  if (A5 > m) { m = A5; jm = 4; }
  const T E = T(12e-6) * (m * m) + tiny_dx2;
  const T t1 = a - b, u1 = a - T(3) * b;
  const T t2 = b - c, u2 = b + c;
  const T t3 = c - d, u3 = T(3) * c - d;
  const T s1 = T(13) * (t1 * t1) + T(3) * (u1 * u1) + E;
  const T s2 = T(13) * (t2 * t2) + T(3) * (u2 * u2) + E;
  const T s3 = T(13) * (t3 * t3) + T(3) * (u3 * u3) + E;
  const T r1 = T(1) / s1, r2 = T(1) / s2, r3 = T(1) / s3;
  const T al1 = T(0.1) * r1 * r1, al2 = T(0.6) * r2 * r2, al3 = T(0.3) * r3 * r3;
  const T rw = T(1) / (al1 + al2 + al3);
  const T w1 = al1 * rw, w2 = al2 * rw, w3 = al3 * rw;
  const T c1 = (a - T(2) * b + c) * T(1.0 / 3.0);
  const T c3 = (b - T(2) * c + d) * T(1.0 / 6.0);
  const T phi = w1 * c1 + w3 * c3;
  F = (T(5) * D3 + T(2) * D4 - D2) * T(1.0 / 6.0) + phi;
  const T sb1 = T(-2) * w1 * (c1 - phi) * r1;
  const T sb2 = T(2) * w2 * phi * r2;
  const T sb3 = T(-2) * w3 * (c3 - phi) * r3;
  const T ab = sb1 * (T(26) * t1 + T(6) * u1) + w1 * T(1.0 / 3.0);
  const T bb = sb1 * (T(-26) * t1 - T(18) * u1) + sb2 * (T(26) * t2 + T(6) * u2) - w1 * T(2.0 / 3.0) +
               w3 * T(1.0 / 6.0);
  const T cb = sb2 * (T(-26) * t2 + T(6) * u2) + sb3 * (T(26) * t3 + T(18) * u3) + w1 * T(1.0 / 3.0) -
               w3 * T(1.0 / 3.0);
  const T db = sb3 * (T(-26) * t3 - T(6) * u3) + w3 * T(1.0 / 6.0);
  g[0] = ab;
  g[1] = bb - ab - T(1.0 / 6.0);
  g[2] = cb - bb + T(5.0 / 6.0);
  g[3] = db - cb + T(1.0 / 3.0);
  g[4] = -db;
  // E through max|D| (the branch taken above)
  const T eb = (sb1 + sb2 + sb3) * T(24e-6) * m;
  const T Dm = jm == 0 ? D1 : jm == 1 ? D2 : jm == 2 ? D3 : jm == 3 ? D4 : D5;
  const T ge = Dm < T(0) ? -eb : eb;
#pragma unroll
  for (int q = 0; q < 5; q++) g[q] += (q == jm) ? ge : T(0);
}

__device__ __forceinline__ int wrap(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

template <typename T>
struct AdjArgs {
  Box fb;
  int n[3];
  Field<const T> ubar, us, v[3];
  Field<T> vbar[3];
  T* slots;       // [axis][q][cell], 18 * ncell
  int64_t ncell;
  T gscale;       // a_s * dt
  T inv[3];       // 1/dx, 1/dy, 1/dz
  T tiny_dx2[3];  // 12 * tiny * dx^2 per axis
  T zero_tol;     // upwind switch threshold, the forward stages' setting
};

template <typename T>
__global__ void __launch_bounds__(256) k_adj_faces(AdjArgs<T> a) {
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncell) return;
  const int nx = a.n[0], ny = a.n[1];
  const int li = (int)(cell % nx);
  const int lj = (int)((cell / nx) % ny);
  const int lk = (int)(cell / ((int64_t)nx * ny));
  const int i = a.fb.ilo + li, j = a.fb.jlo + lj, k = a.fb.klo + lk;
  const T g = a.gscale * a.ubar.at(i, j, k);
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    const T vel = a.v[ax].at(i, j, k);
    T slot[6];
#pragma unroll
    for (int q = 0; q < 6; q++) slot[q] = T(0);
    if (fabs(vel) >= a.zero_tol) {
      T p[7];
      const int l0 = ax == 0 ? li : ax == 1 ? lj : lk;
      const int n = a.n[ax];
#pragma unroll
      for (int o = 0; o < 7; o++) {
        const int l = wrap(l0 + o - 3, n);
        const int ii = ax == 0 ? a.fb.ilo + l : i;
        const int jj = ax == 1 ? a.fb.jlo + l : j;
        const int kk = ax == 2 ? a.fb.klo + l : k;
        p[o] = a.us.at(ii, jj, kk);
      }
      T d[6];
#pragma unroll
      for (int q = 0; q < 6; q++) d[q] = p[q + 1] - p[q];
      const bool bwd = vel > T(0);
      T F, gr[5];
      if (bwd) weno5_grad<T>(d[0], d[1], d[2], d[3], d[4], a.tiny_dx2[ax], F, gr);
      else weno5_grad<T>(d[5], d[4], d[3], d[2], d[1], a.tiny_dx2[ax], F, gr);
      const T coef = -g * vel * a.inv[ax];
      if (bwd) {
#pragma unroll
        for (int q = 0; q < 5; q++) slot[q] = coef * gr[q];
      } else {
#pragma unroll
        for (int q = 0; q < 5; q++) slot[5 - q] = coef * gr[q];
      }
      if (a.vbar[ax].p) a.vbar[ax].at(i, j, k) -= g * (F * a.inv[ax]);
    }
#pragma unroll
    for (int q = 0; q < 6; q++) a.slots[(int64_t)(ax * 6 + q) * a.ncell + cell] = slot[q];
  }
}

template <typename T>
struct GatherArgs {
  Box fb;
  int n[3];
  Field<const T> ubar;
  Field<T> adj_s, adj_c;
  const T* slots;
  int64_t ncell;
  T as, ac;
  int acc_s, acc_c;
};

template <typename T>
__global__ void __launch_bounds__(256) k_adj_gather(GatherArgs<T> a) {
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncell) return;
  const int nx = a.n[0], ny = a.n[1];
  const int li = (int)(cell % nx);
  const int lj = (int)((cell / nx) % ny);
  const int lk = (int)(cell / ((int64_t)nx * ny));
  const int i = a.fb.ilo + li, j = a.fb.jlo + lj, k = a.fb.klo + lk;
  const int64_t stride[3] = {1, (int64_t)nx, (int64_t)nx * ny};
  const int l[3] = {li, lj, lk};
  T sum = T(0);
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    const int n = a.n[ax];
    const T* s = a.slots + (int64_t)(ax * 6) * a.ncell + cell - (int64_t)l[ax] * stride[ax];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int64_t lo = (int64_t)wrap(l[ax] + 2 - q, n) * stride[ax];
      const int64_t hi = (int64_t)wrap(l[ax] + 3 - q, n) * stride[ax];
      sum += s[(int64_t)q * a.ncell + lo] - s[(int64_t)q * a.ncell + hi];
    }
  }
  const T ub = a.ubar.at(i, j, k);
  const T vs = a.as * ub + sum;
  T& o = a.adj_s.at(i, j, k);
  o = a.acc_s ? o + vs : vs;
  if (a.adj_c.p) {
    T& c = a.adj_c.at(i, j, k);
    c = a.acc_c ? c + a.ac * ub : a.ac * ub;
  }
}

// ---------------------------------------------------------------------------
// Heaviside mismatch: loss = sum over the fill box of (H(phi) - H(target))^2
// with the smoothed Heaviside of lsm_calculus_toolbox.f:25-54
//   H(x) = 0 (x < -eps), 1 (x > eps), 0.5 (1 + x/eps + one_over_pi sin(pi x/eps))
// accumulated in FP64; grad (optional) = d loss / d phi.
template <typename T>
__device__ __forceinline__ void heaviside(T x, T eps, T& H, T& dH) {
  const T pi = T(3.14159265358979323846), one_over_pi = T(0.31830988618379);
  if (x < -eps) {
    H = T(0);
    dH = T(0);
  } else if (x > eps) {
    H = T(1);
    dH = T(0);
  } else {
    const T ioe = T(1) / eps;
    H = T(0.5) * (T(1) + x * ioe + one_over_pi * sin(pi * x * ioe));
    dH = T(0.5) * (ioe + one_over_pi * cos(pi * x * ioe) * pi * ioe);
  }
}

constexpr int kLossBlocks = 1024;

template <typename T>
__global__ void __launch_bounds__(256)
    k_heaviside_mismatch(Box fb, Field<const T> phi, Field<const T> tgt, T eps, Field<T> grad, double* part) {
  const int nx = fb.nx(), ny = fb.ny();
  const int64_t ncell = fb.cells();
  double acc = 0.0;
  for (int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
       cell += (int64_t)gridDim.x * blockDim.x) {
    const int i = fb.ilo + (int)(cell % nx);
    const int j = fb.jlo + (int)((cell / nx) % ny);
    const int k = fb.klo + (int)(cell / ((int64_t)nx * ny));
    T h, dh, ht, dht;
    heaviside<T>(phi.at(i, j, k), eps, h, dh);
    heaviside<T>(tgt.at(i, j, k), eps, ht, dht);
    const T diff = h - ht;
    acc += (double)diff * (double)diff;
    if (grad.p) grad.at(i, j, k) = T(2) * diff * dh;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* part, int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int q = (int)threadIdx.x; q < n; q += 256) acc += part[q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

namespace {
typedef real_t R;

Box box6(const int* b) { return Box{b[0], b[1], b[2], b[3], b[4], b[5]}; }

bool dev_only(const char* fn, const void* p, const char* what) {
  if (p && !is_device_pointer(p)) {
    raise_error(std::string(fn) + ": " + what + " must be device memory");
    return false;
  }
  return true;
}

template <typename T>
Field<T> field_or_null(T* p, const int* gb) {
  if (!p) {
    Field<T> f;
    f.p = nullptr;
    f.ilo = f.jlo = f.klo = 0;
    f.sy = f.sz = 0;
    return f;
  }
  return make_field<T>(p, box6(gb));
}

bool rk_weights(int form, R& as, R& ac) {
  switch (form) {
    case RK_EULER: as = R(1); ac = R(0); return true;
    case RK2_STAGE2: as = R(0.5); ac = R(0.5); return true;
    case RK3_STAGE2: as = R(0.25); ac = R(0.75); return true;
    case RK3_STAGE3: as = R(2.0 / 3.0); ac = R(1.0 / 3.0); return true;
    default: return false;
  }
}
}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" {

long long lsm3d_amd_advect_adjoint_work_elems(const int fb[6]) { return 18LL * box6(fb).cells(); }

void lsm3d_amd_advect_stage_adjoint(int rk_form, const LSMLIB_REAL* adj_out, const int adj_out_gb[6],
                                    LSMLIB_REAL* adj_stage, const int adj_stage_gb[6], LSMLIB_REAL* adj_cur,
                                    const int adj_cur_gb[6], int accumulate, const LSMLIB_REAL* phi_stage,
                                    const int stage_gb[6], const LSMLIB_REAL* vel_x, const LSMLIB_REAL* vel_y,
                                    const LSMLIB_REAL* vel_z, const int vel_gb[6], LSMLIB_REAL* vel_adj_x,
                                    LSMLIB_REAL* vel_adj_y, LSMLIB_REAL* vel_adj_z, const int vel_adj_gb[6],
                                    LSMLIB_REAL* work, const int fbp[6], LSMLIB_REAL dx, LSMLIB_REAL dy,
                                    LSMLIB_REAL dz, LSMLIB_REAL dt, int periodic_mask) {
  const char* fn = "lsm3d_amd_advect_stage_adjoint";
  Call c(fn);
  const Box fb = box6(fbp);
  R as = 0, ac = 0;
  if (!rk_weights(rk_form, as, ac)) raise_error(std::string(fn) + ": unknown Runge-Kutta form");
  if ((periodic_mask & 7) != 7) raise_error(std::string(fn) + ": only boxes periodic on every axis");
  if (fb.nx() < 6 || fb.ny() < 6 || fb.nz() < 6) raise_error(std::string(fn) + ": periodic axes need 6 cells");
  if (!adj_out || !adj_stage || !phi_stage || !vel_x || !vel_y || !vel_z || !work)
    raise_error(std::string(fn) + ": missing array");
  if (!c.ok()) return;
  require_inside(fn, "adj_out", box6(adj_out_gb), fb, 0, 0, 0);
  require_inside(fn, "adj_stage", box6(adj_stage_gb), fb, 0, 0, 0);
  if (adj_cur) require_inside(fn, "adj_cur", box6(adj_cur_gb), fb, 0, 0, 0);
  require_inside(fn, "phi_stage", box6(stage_gb), fb, 0, 0, 0);
  require_inside(fn, "vel", box6(vel_gb), fb, 0, 0, 0);
  const bool with_vbar = vel_adj_x || vel_adj_y || vel_adj_z;
  if (with_vbar) {
    if (!(vel_adj_x && vel_adj_y && vel_adj_z)) raise_error(std::string(fn) + ": give all three vel_adj arrays");
    else require_inside(fn, "vel_adj", box6(vel_adj_gb), fb, 0, 0, 0);
  }
  if (!(dev_only(fn, adj_out, "adj_out") && dev_only(fn, adj_stage, "adj_stage") &&
        dev_only(fn, adj_cur, "adj_cur") && dev_only(fn, phi_stage, "phi_stage") &&
        dev_only(fn, vel_x, "vel_x") && dev_only(fn, vel_y, "vel_y") && dev_only(fn, vel_z, "vel_z") &&
        dev_only(fn, vel_adj_x, "vel_adj_x") && dev_only(fn, vel_adj_y, "vel_adj_y") &&
        dev_only(fn, vel_adj_z, "vel_adj_z") && dev_only(fn, work, "work")))
    return;
  if (!c.ok()) return;
  const int64_t ncell = fb.cells();
  const unsigned nblk = (unsigned)((ncell + 255) / 256);
  if ((ncell + 255) / 256 > 0x7fffffffLL) {
    raise_error(std::string(fn) + ": box too large for one launch");
    return;
  }
  AdjArgs<R> a;
  a.fb = fb;
  a.n[0] = fb.nx();
  a.n[1] = fb.ny();
  a.n[2] = fb.nz();
  a.ubar = make_field<const R>(adj_out, box6(adj_out_gb));
  a.us = make_field<const R>(phi_stage, box6(stage_gb));
  a.v[0] = make_field<const R>(vel_x, box6(vel_gb));
  a.v[1] = make_field<const R>(vel_y, box6(vel_gb));
  a.v[2] = make_field<const R>(vel_z, box6(vel_gb));
  a.vbar[0] = field_or_null<R>(vel_adj_x, vel_adj_gb);
  a.vbar[1] = field_or_null<R>(vel_adj_y, vel_adj_gb);
  a.vbar[2] = field_or_null<R>(vel_adj_z, vel_adj_gb);
  a.slots = work;
  a.ncell = ncell;
  a.gscale = as * dt;
  a.zero_tol = (R)current_zero_tol();  // the switch the forward stage used (lsm3d_amd_set_zero_tol)
  const R h[3] = {dx, dy, dz};
  for (int d = 0; d < 3; d++) {
    a.inv[d] = R(1) / h[d];
    a.tiny_dx2[d] = R(12) * Tol<R>::tiny * h[d] * h[d];
  }
  hipLaunchKernelGGL(k_adj_faces<R>, dim3(nblk), dim3(256), 0, c.stream(), a);
  c.launched("k_adj_faces");
  GatherArgs<R> gth;
  gth.fb = fb;
  gth.n[0] = a.n[0];
  gth.n[1] = a.n[1];
  gth.n[2] = a.n[2];
  gth.ubar = a.ubar;
  gth.adj_s = make_field<R>(adj_stage, box6(adj_stage_gb));
  gth.adj_c = field_or_null<R>(adj_cur, adj_cur_gb);
  gth.slots = work;
  gth.ncell = ncell;
  gth.as = as;
  gth.ac = ac;
  gth.acc_s = accumulate & 1;
  gth.acc_c = (accumulate >> 1) & 1;
  hipLaunchKernelGGL(k_adj_gather<R>, dim3(nblk), dim3(256), 0, c.stream(), gth);
  c.launched("k_adj_gather");
  c.finish();
}

void lsm3d_amd_heaviside_mismatch(const LSMLIB_REAL* phi, const int phi_gb[6], const LSMLIB_REAL* target,
                                  const int target_gb[6], const int fbp[6], LSMLIB_REAL epsilon, double* loss,
                                  LSMLIB_REAL* grad, const int grad_gb[6]) {
  const char* fn = "lsm3d_amd_heaviside_mismatch";
  Call c(fn);
  const Box fb = box6(fbp);
  if (!phi || !target || !loss) raise_error(std::string(fn) + ": missing array");
  if (!c.ok()) return;
  require_inside(fn, "phi", box6(phi_gb), fb, 0, 0, 0);
  require_inside(fn, "target", box6(target_gb), fb, 0, 0, 0);
  if (grad) require_inside(fn, "grad", box6(grad_gb), fb, 0, 0, 0);
  if (!(dev_only(fn, phi, "phi") && dev_only(fn, target, "target") && dev_only(fn, loss, "loss") &&
        dev_only(fn, grad, "grad")))
    return;
  if (!c.ok()) return;
  double* part = static_cast<double*>(device_scratch(kLossBlocks * sizeof(double)));
  if (!part) return;
  hipLaunchKernelGGL(k_heaviside_mismatch<R>, dim3(kLossBlocks), dim3(256), 0, c.stream(), fb,
                     make_field<const R>(phi, box6(phi_gb)), make_field<const R>(target, box6(target_gb)),
                     epsilon, field_or_null<R>(grad, grad_gb), part);
  c.launched("k_heaviside_mismatch");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c.stream(), part, kLossBlocks, loss);
  c.launched("k_sum_partials");
  c.finish();
}

}  // extern "C"
