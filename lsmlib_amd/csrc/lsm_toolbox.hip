// gfx950 kernels behind LSMLIB's LSM3D_* toolbox entry points, and the
// extern "C" Fortran-77-convention entry points themselves.
//
// Each entry point keeps the reference prototype (every argument by
// reference, one ghost box per array, a fill/include box, spacings), checks
// that every stencil read stays inside its array's ghost box, accepts device
// or host memory (host arrays are staged through device buffers), queues one
// kernel on the library stream and, like the Fortran, completes before it
// returns (unless lsm_amd_set_sync(0) asked for stream ordering only).
//
// The D1/D2/D3 scratch arrays of the derivative routines are accepted and
// not written: no caller reads them after a call, and the kernels form the
// undivided differences in registers instead of re-reading them from HBM.
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

using lsm::Box;
using lsm::Call;
using lsm::Field;

namespace lsm {

// ============================================================================
// Spatial derivatives
// ============================================================================
template <typename T, int S, int DIR>
__device__ __forceinline__ void gather(const Field<const T>& phi, int i, int j, int k, T* p) {
  constexpr int R = SchemeRadius<S>::R;
#pragma unroll
  for (int q = -R; q <= R; q++)
    p[3 + q] = phi.at(i + (DIR == 0 ? q : 0), j + (DIR == 1 ? q : 0), k + (DIR == 2 ? q : 0));
}

template <typename T, int S>
__global__ void __launch_bounds__(256)
    k_hj_pm(Field<T> xp, Field<T> yp, Field<T> zp, Field<T> xm, Field<T> ym, Field<T> zm,
            Field<const T> phi, Box fb, T ix, T iy, T iz) {
  LSM_CELL_LOOP_BEGIN(fb)
  T p[7];
  gather<T, S, 0>(phi, i, j, k, p);
  xp.at(i, j, k) = hj_plus<T, S>(p, ix);
  xm.at(i, j, k) = hj_minus<T, S>(p, ix);
  gather<T, S, 1>(phi, i, j, k, p);
  yp.at(i, j, k) = hj_plus<T, S>(p, iy);
  ym.at(i, j, k) = hj_minus<T, S>(p, iy);
  gather<T, S, 2>(phi, i, j, k, p);
  zp.at(i, j, k) = hj_plus<T, S>(p, iz);
  zm.at(i, j, k) = hj_minus<T, S>(p, iz);
  LSM_CELL_LOOP_END
}

template <typename T, int S>
__global__ void __launch_bounds__(256)
    k_hj_upwind(Field<T> gx, Field<T> gy, Field<T> gz, Field<const T> phi, Field<const T> vx,
                Field<const T> vy, Field<const T> vz, Box fb, T ix, T iy, T iz) {
  LSM_CELL_LOOP_BEGIN(fb)
  T p[7];
  gather<T, S, 0>(phi, i, j, k, p);
  gx.at(i, j, k) = hj_upwind<T, S>(p, vx.at(i, j, k), ix);
  gather<T, S, 1>(phi, i, j, k, p);
  gy.at(i, j, k) = hj_upwind<T, S>(p, vy.at(i, j, k), iy);
  gather<T, S, 2>(phi, i, j, k, p);
  gz.at(i, j, k) = hj_upwind<T, S>(p, vz.at(i, j, k), iz);
  LSM_CELL_LOOP_END
}

// ============================================================================
// Central differences  (lsm_spatial_derivatives3d.f.in:3016-3264, 3586-3650)
// Products and sums are rounded one by one, as the reference's are.
// ============================================================================
// :3016-3088
template <typename T>
__global__ void __launch_bounds__(256)
    k_central_grad2(Field<T> gx, Field<T> gy, Field<T> gz, Field<const T> phi, Box fb, T fx, T fy, T fz) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  gx.at(i, j, k) = (phi.at(i + 1, j, k) - phi.at(i - 1, j, k)) * fx;
  gy.at(i, j, k) = (phi.at(i, j + 1, k) - phi.at(i, j - 1, k)) * fy;
  gz.at(i, j, k) = (phi.at(i, j, k + 1) - phi.at(i, j, k - 1)) * fz;
  LSM_CELL_LOOP_END
}

// :3094-3174
template <typename T>
__global__ void __launch_bounds__(256)
    k_central_grad4(Field<T> gx, Field<T> gy, Field<T> gz, Field<const T> phi, Box fb, T fx, T fy, T fz) {
#pragma clang fp contract(off)
  const T e = T(8);
  LSM_CELL_LOOP_BEGIN(fb)
  gx.at(i, j, k) = (((-phi.at(i + 2, j, k) + e * phi.at(i + 1, j, k)) + phi.at(i - 2, j, k)) -
                    e * phi.at(i - 1, j, k)) * fx;
  gy.at(i, j, k) = (((-phi.at(i, j + 2, k) + e * phi.at(i, j + 1, k)) + phi.at(i, j - 2, k)) -
                    e * phi.at(i, j - 1, k)) * fy;
  gz.at(i, j, k) = (((-phi.at(i, j, k + 2) + e * phi.at(i, j, k + 1)) + phi.at(i, j, k - 2)) -
                    e * phi.at(i, j, k - 1)) * fz;
  LSM_CELL_LOOP_END
}

// :3180-3264
template <typename T>
__global__ void __launch_bounds__(256)
    k_laplacian2(Field<T> lap, Field<const T> phi, Box fb, T ix2, T iy2, T iz2) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T c = T(2) * phi.at(i, j, k);
  lap.at(i, j, k) = (ix2 * ((phi.at(i + 1, j, k) - c) + phi.at(i - 1, j, k)) +
                     iy2 * ((phi.at(i, j + 1, k) - c) + phi.at(i, j - 1, k))) +
                    iz2 * ((phi.at(i, j, k + 1) - c) + phi.at(i, j, k - 1));
  LSM_CELL_LOOP_END
}

// First-order delta function, lsm_calculus_toolbox3d.f.in:15-120: at each
// cell the sum over the six neighbours across which phi changes sign (or
// touches zero) of |phi_nb n / (dx (phi_nb - phi))|.  The z terms use the x
// component of the normal and 1/dx, as the reference writes them (:98-110).
template <typename T>
__global__ void __launch_bounds__(256)
    k_delta_order1(Field<T> delta, Field<const T> phi, Field<const T> nx, Field<const T> ny, Box fb, T odx,
                   T ody) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T c = phi.at(i, j, k), ax = nx.at(i, j, k), ay = ny.at(i, j, k);
  const T nb[6] = {phi.at(i + 1, j, k), phi.at(i - 1, j, k), phi.at(i, j + 1, k),
                   phi.at(i, j - 1, k), phi.at(i, j, k + 1), phi.at(i, j, k - 1)};
  T sum = T(0);
#pragma unroll
  for (int q = 0; q < 6; q++) {
    T t = T(0);
    if (nb[q] * c <= T(0)) {
      t = nb[q] * (q == 2 || q == 3 ? ay : ax) * (q == 2 || q == 3 ? ody : odx);
      t = t / (nb[q] - c);
    }
    sum = (q == 0) ? fabs(t) : sum + fabs(t);
  }
  delta.at(i, j, k) = sum;
  LSM_CELL_LOOP_END
}

// Mean / Gaussian curvature, lsm_curvature3d.f.in (one thread per cell)
template <typename T, int ORDER, bool GAUSS>
__global__ void __launch_bounds__(256)
    k_curvature(Field<T> kappa, Field<const T> phi, Field<const T> px, Field<const T> py, Field<const T> pz,
                Field<const T> gm, Box fb, CurvCoef<T> c) {
  LSM_CELL_LOOP_BEGIN(fb)
  kappa.at(i, j, k) = curvature_at<T, ORDER, GAUSS>(phi, px, py, pz, gm, i, j, k, c);
  LSM_CELL_LOOP_END
}

// dst = src over the fill box, lsm_samrai_f77_utilities.f:159-169
template <typename T>
__global__ void __launch_bounds__(256) k_copy_box(Field<T> dst, Field<const T> src, Box fb) {
  LSM_CELL_LOOP_BEGIN(fb)
  dst.at(i, j, k) = src.at(i, j, k);
  LSM_CELL_LOOP_END
}

// rhs += b kappa |grad phi|, lsm_level_set_evolution3d.f.in:513-521
template <typename T>
__global__ void __launch_bounds__(256)
    k_precomputed_curv(Field<T> rhs, Field<const T> kappa, Field<const T> gm, Box fb, T b) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  rhs.at(i, j, k) = rhs.at(i, j, k) + (b * kappa.at(i, j, k)) * gm.at(i, j, k);
  LSM_CELL_LOOP_END
}

// :3586-3650 (tol = 1e-8 on the squares and on the magnitude)
template <typename T>
__global__ void __launch_bounds__(256)
    k_grad_mag(Field<T> mag, Field<const T> gx, Field<const T> gy, Field<const T> gz, Box fb) {
#pragma clang fp contract(off)
  const T tol = T(1e-8), z = T(0);
  LSM_CELL_LOOP_BEGIN(fb)
  const T a = gx.at(i, j, k), b = gy.at(i, j, k), c = gz.at(i, j, k);
  T x2 = a * a, y2 = b * b, z2 = c * c;
  if (x2 < tol) x2 = z;
  if (y2 < tol) y2 = z;
  if (z2 < tol) z2 = z;
  const T t = sqrt((x2 + y2) + z2);
  mag.at(i, j, k) = (t < tol) ? z : t;
  LSM_CELL_LOOP_END
}

// ============================================================================
// Normals and the field-extension right-hand side
// ============================================================================
// phi-upwind choice of one component, lsm_spatial_derivatives3d.f.in:3270-3480:
// false (F left unchanged) when phi == 0.
template <typename T>
__device__ __forceinline__ bool phi_upwind_pick(T pc, T pm, T pp, T fp, T fm, T& out) {
  if (pc > T(0)) {
    if (pm <= pc) out = (pp < pm) ? fp : fm;
    else if (pp <= pc) out = fp;
    else out = T(0);
    return true;
  }
  if (pc < T(0)) {
    if (pm >= pc) out = (pp > pm) ? fp : fm;
    else if (pp >= pc) out = fp;
    else out = T(0);
    return true;
  }
  return false;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_phi_upwind_grad(Field<T> fx, Field<T> fy, Field<T> fz, Field<const T> fxp, Field<const T> fyp,
                      Field<const T> fzp, Field<const T> fxm, Field<const T> fym, Field<const T> fzm,
                      Field<const T> phi, Box fb) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T pc = phi.at(i, j, k);
  T v;
  if (phi_upwind_pick<T>(pc, phi.at(i - 1, j, k), phi.at(i + 1, j, k), fxp.at(i, j, k), fxm.at(i, j, k), v))
    fx.at(i, j, k) = v;
  if (phi_upwind_pick<T>(pc, phi.at(i, j - 1, k), phi.at(i, j + 1, k), fyp.at(i, j, k), fym.at(i, j, k), v))
    fy.at(i, j, k) = v;
  if (phi_upwind_pick<T>(pc, phi.at(i, j, k - 1), phi.at(i, j, k + 1), fzp.at(i, j, k), fzm.at(i, j, k), v))
    fz.at(i, j, k) = v;
  LSM_CELL_LOOP_END
}

// lsm_spatial_derivatives3d.f.in:3486-3580
template <typename T>
__global__ void __launch_bounds__(256)
    k_average_grad(Field<T> gx, Field<T> gy, Field<T> gz, Field<const T> xp, Field<const T> yp, Field<const T> zp,
                   Field<const T> xm, Field<const T> ym, Field<const T> zm, Box fb) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  gx.at(i, j, k) = T(0.5) * (xp.at(i, j, k) + xm.at(i, j, k));
  gy.at(i, j, k) = T(0.5) * (yp.at(i, j, k) + ym.at(i, j, k));
  gz.at(i, j, k) = T(0.5) * (zp.at(i, j, k) + zm.at(i, j, k));
  LSM_CELL_LOOP_END
}

// lsm_geometry3d_fort.f.in:27-110 (unsigned) and :129-240 (signed: sgn(phi)
// smoothed with max(dx)^2, zero where |phi| <= zero_tol)
template <typename T, bool SIGNED>
__global__ void __launch_bounds__(256)
    k_unit_normal(Field<T> nx, Field<T> ny, Field<T> nz, Field<const T> px, Field<const T> py, Field<const T> pz,
                  Field<const T> phi, Box fb, T dx_sq) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T x = px.at(i, j, k), y = py.at(i, j, k), z = pz.at(i, j, k);
  T a = T(1), b = T(0), c = T(0);
  if (SIGNED) {
    const T pc = phi.at(i, j, k);
    if (fabs(pc) > Tol<T>::zero) {
      const T n2 = (x * x + y * y) + z * z;
      if (n2 >= Tol<T>::zero) {
        const T sgn = pc / sqrt(pc * pc + n2 * dx_sq);
        const T inv = T(1) / sqrt(n2);
        a = (sgn * x) * inv;
        b = (sgn * y) * inv;
        c = (sgn * z) * inv;
      }
    } else {
      a = T(0);
    }
  } else {
    const T norm = sqrt((x * x + y * y) + z * z);
    if (norm >= Tol<T>::zero) {
      const T inv = T(1) / norm;
      a = x * inv;
      b = y * inv;
      c = z * inv;
    }
  }
  nx.at(i, j, k) = a;
  ny.at(i, j, k) = b;
  nz.at(i, j, k) = c;
  LSM_CELL_LOOP_END
}

// lsm_field_extension3d.f:31-158: -(n . grad S) unless |phi| <= cutoff and
// phi changes sign towards one of the six neighbours (then 0)
template <typename T>
__device__ __forceinline__ bool field_ext_active(const Field<const T>& phi, int i, int j, int k, T cutoff) {
#pragma clang fp contract(off)
  const T pc = phi.at(i, j, k), z = T(0);
  if (fabs(pc) > cutoff) return true;
  return pc * phi.at(i - 1, j, k) > z && pc * phi.at(i + 1, j, k) > z && pc * phi.at(i, j - 1, k) > z &&
         pc * phi.at(i, j + 1, k) > z && pc * phi.at(i, j, k - 1) > z && pc * phi.at(i, j, k + 1) > z;
}

// The field-extension velocity: the signed normal where the extension
// equation is active, 0 where LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS sets the
// right-hand side to zero.  With it, -V.grad(S) (upwind by V) equals that
// routine's right-hand side cell by cell, so the fused advection stages
// advance the extension equation.
template <typename T>
__global__ void __launch_bounds__(256)
    k_field_ext_velocity(Field<T> vx, Field<T> vy, Field<T> vz, Field<const T> nx, Field<const T> ny,
                         Field<const T> nz, Field<const T> phi, Box fb, T cutoff) {
  LSM_CELL_LOOP_BEGIN(fb)
  const bool on = field_ext_active<T>(phi, i, j, k, cutoff);
  vx.at(i, j, k) = on ? nx.at(i, j, k) : T(0);
  vy.at(i, j, k) = on ? ny.at(i, j, k) : T(0);
  vz.at(i, j, k) = on ? nz.at(i, j, k) : T(0);
  LSM_CELL_LOOP_END
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_field_ext_rhs(Field<T> rhs, Field<const T> phi, Field<const T> sx, Field<const T> sy, Field<const T> sz,
                    Field<const T> nx, Field<const T> ny, Field<const T> nz, Box fb, T cutoff) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  T r = T(0);
  if (field_ext_active<T>(phi, i, j, k, cutoff))
    r = -((nx.at(i, j, k) * sx.at(i, j, k) + ny.at(i, j, k) * sy.at(i, j, k)) + nz.at(i, j, k) * sz.at(i, j, k));
  rhs.at(i, j, k) = r;
  LSM_CELL_LOOP_END
}

// ============================================================================
// Level set equation RHS  (lsm_level_set_evolution3d.f.in)
// ============================================================================
template <typename T>
__global__ void __launch_bounds__(256) k_fill(Field<T> a, Box b, T v) {
  LSM_CELL_LOOP_BEGIN(b)
  a.at(i, j, k) = v;
  LSM_CELL_LOOP_END
}

// :113-116
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_advection(Field<T> rhs, Field<const T> px, Field<const T> py, Field<const T> pz,
                    Field<const T> vx, Field<const T> vy, Field<const T> vz, Box fb) {
  LSM_CELL_LOOP_BEGIN(fb)
  rhs.at(i, j, k) = rhs.at(i, j, k) - (vx.at(i, j, k) * px.at(i, j, k) + vy.at(i, j, k) * py.at(i, j, k) +
                                       vz.at(i, j, k) * pz.at(i, j, k));
  LSM_CELL_LOOP_END
}

// :131-242 (field vn) and :248-350 (constant vn: vn_field.p == nullptr)
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_normal_vel(Field<T> rhs, Field<const T> xp, Field<const T> yp, Field<const T> zp,
                     Field<const T> xm, Field<const T> ym, Field<const T> zm, Field<const T> vnf,
                     T vn_const, Box fb) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T vn = vnf.p ? vnf.at(i, j, k) : vn_const;
  if (fabs(vn) >= Tol<T>::zero) {
    const T n2 = godunov_norm_sq<T>(vn, xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k), xm.at(i, j, k),
                                    ym.at(i, j, k), zm.at(i, j, k));
    rhs.at(i, j, k) = rhs.at(i, j, k) - vn * sqrt(n2);
  }
  LSM_CELL_LOOP_END
}

// :356-460
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_const_curv(Field<T> rhs, Field<const T> px, Field<const T> py, Field<const T> pz, Field<const T> pxx,
                     Field<const T> pxy, Field<const T> pxz, Field<const T> pyy, Field<const T> pyz,
                     Field<const T> pzz, Box fb, T b) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T curv = curvature_term<T>(px.at(i, j, k), py.at(i, j, k), pz.at(i, j, k), pxx.at(i, j, k),
                                   pxy.at(i, j, k), pxz.at(i, j, k), pyy.at(i, j, k), pyz.at(i, j, k),
                                   pzz.at(i, j, k), Tol<T>::zero);
  rhs.at(i, j, k) = rhs.at(i, j, k) + b * curv;
  LSM_CELL_LOOP_END
}

// Lax-Friedrichs advection term (lf_axis): rhs -= (H_x + H_y) + H_z, alpha
// per axis = |v_a| at the cell (mode 0) or the given global alpha[a].
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_lf_advection(Field<T> rhs, Field<const T> xp, Field<const T> yp, Field<const T> zp,
                       Field<const T> xm, Field<const T> ym, Field<const T> zm, Field<const T> vx,
                       Field<const T> vy, Field<const T> vz, Box fb, int mode, T ax, T ay, T az) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T u = vx.at(i, j, k), v = vy.at(i, j, k), w = vz.at(i, j, k);
  const T hx = lf_axis<T>(u, mode ? ax : fabs(u), xp.at(i, j, k), xm.at(i, j, k));
  const T hy = lf_axis<T>(v, mode ? ay : fabs(v), yp.at(i, j, k), ym.at(i, j, k));
  const T hz = lf_axis<T>(w, mode ? az : fabs(w), zp.at(i, j, k), zm.at(i, j, k));
  rhs.at(i, j, k) = rhs.at(i, j, k) - ((hx + hy) + hz);
  LSM_CELL_LOOP_END
}

// Lax-Friedrichs normal-speed term (lf_normal): rhs -= H, alpha = |vn| at the
// cell (mode 0) or the given global alpha; vn from the field, or vn_const
// when vnf.p is null.
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_lf_normal(Field<T> rhs, Field<const T> xp, Field<const T> yp, Field<const T> zp, Field<const T> xm,
                    Field<const T> ym, Field<const T> zm, Field<const T> vnf, T vn_const, Box fb, int mode,
                    T alpha) {
#pragma clang fp contract(off)
  LSM_CELL_LOOP_BEGIN(fb)
  const T vn = vnf.p ? vnf.at(i, j, k) : vn_const;
  const T h = lf_normal<T>(vn, mode ? alpha : fabs(vn), xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k),
                           xm.at(i, j, k), ym.at(i, j, k), zm.at(i, j, k));
  rhs.at(i, j, k) = rhs.at(i, j, k) - h;
  LSM_CELL_LOOP_END
}

// :532-719 over the reference's loop box (see the entry point)
template <typename T>
__global__ void __launch_bounds__(256)
    k_add_ext_normal_vel(Field<T> rhs, Field<const T> xp, Field<const T> yp, Field<const T> zp,
                         Field<const T> xm, Field<const T> ym, Field<const T> zm, Field<const T> vn,
                         Field<const T> vx, Field<const T> vy, Field<const T> vz, Box lb) {
  LSM_CELL_LOOP_BEGIN(lb)
  const T vnc = vn.at(i, j, k), vxc = vx.at(i, j, k), vyc = vy.at(i, j, k), vzc = vz.at(i, j, k);
  const T px = ext_normal_select<T>(vxc, vnc, xp.at(i, j, k), xm.at(i, j, k));
  const T py = ext_normal_select<T>(vyc, vnc, yp.at(i, j, k), ym.at(i, j, k));
  const T pz = ext_normal_select<T>(vzc, vnc, zp.at(i, j, k), zm.at(i, j, k));
  const T n2 = px * px + py * py + pz * pz;
  T r = rhs.at(i, j, k) - vxc * px - vyc * py - vzc * pz;
  if (fabs(vnc) >= Tol<T>::zero && n2 >= Tol<T>::zero) r = r - vnc * sqrt(n2);
  rhs.at(i, j, k) = r;
  LSM_CELL_LOOP_END
}

// ============================================================================
// TVD Runge-Kutta stages  (lsm_tvd_runge_kutta3d.f)
// ============================================================================
template <typename T>
__global__ void __launch_bounds__(256)
    k_rk(int form, Field<T> out, Field<const T> us, Field<const T> uc, Field<const T> rhs, Box fb, T dt) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T ucv = uc.at(i, j, k);
  const T usv = us.p ? us.at(i, j, k) : ucv;
  out.at(i, j, k) = rk_combine<T>(form, ucv, usv, rhs.at(i, j, k), dt);
  LSM_CELL_LOOP_END
}

// ============================================================================
// Reinitialization / orthogonalization  (lsm_reinitialization3d.f.in)
// ============================================================================
// :48-268
template <typename T>
__global__ void __launch_bounds__(256)
    k_reinit_rhs(Field<T> rhs, Field<const T> phi, Field<const T> phi0, Field<const T> xp,
                 Field<const T> yp, Field<const T> zp, Field<const T> xm, Field<const T> ym,
                 Field<const T> zm, Box fb, T dx_sq, int use_phi0) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T cur = use_phi0 ? phi0.at(i, j, k) : phi.at(i, j, k);
  T r = T(0);
  if (fabs(cur) >= Tol<T>::zero) {
    const T n2 = godunov_star_sq<T>(cur, xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k),
                                    xm.at(i, j, k), ym.at(i, j, k), zm.at(i, j, k));
    const T sgn = use_phi0 ? cur / sqrt(cur * cur + dx_sq) : cur / sqrt(cur * cur + n2 * dx_sq);
    r = sgn * (T(1) - sqrt(n2));
  }
  rhs.at(i, j, k) = r;
  LSM_CELL_LOOP_END
}

// :296-471
template <typename T>
__global__ void __launch_bounds__(256)
    k_ortho_rhs(Field<T> rhs, Field<const T> xp, Field<const T> yp, Field<const T> zp,
                Field<const T> xm, Field<const T> ym, Field<const T> zm, Field<const T> psi,
                Field<const T> sxp, Field<const T> syp, Field<const T> szp, Field<const T> sxm,
                Field<const T> sym, Field<const T> szm, Box fb, T dx_sq) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T g0 = T(0.5) * (sxp.at(i, j, k) + sxm.at(i, j, k));
  const T g1 = T(0.5) * (syp.at(i, j, k) + sym.at(i, j, k));
  const T g2 = T(0.5) * (szp.at(i, j, k) + szm.at(i, j, k));
  const T ng = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
  const T ps = psi.at(i, j, k);
  T r = T(0);
  if (fabs(ps) > Tol<T>::zero && ng > Tol<T>::zero) {
    const T s = ps / sqrt(ps * ps + dx_sq * (ng * ng));
    const T z = T(0);
    T acc = fmax(s * g0, z) * xm.at(i, j, k);
    acc = acc + fmin(s * g0, z) * xp.at(i, j, k);
    acc = acc + fmax(s * g1, z) * ym.at(i, j, k);
    acc = acc + fmin(s * g1, z) * yp.at(i, j, k);
    acc = acc + fmax(s * g2, z) * zm.at(i, j, k);
    acc = acc + fmin(s * g2, z) * zp.at(i, j, k);
    r = T(-1) / ng * acc;
  }
  rhs.at(i, j, k) = r;
  LSM_CELL_LOOP_END
}

// :477-581 (Russo-Smereka distance estimate near the interface)
template <typename T>
__global__ void __launch_bounds__(256)
    k_subcell_distance(Field<T> dist0, Field<const T> phi0, Box fb, T flag) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T c = phi0.at(i, j, k);
  const T z = T(0);
  const T nb[6] = {phi0.at(i + 1, j, k), phi0.at(i - 1, j, k), phi0.at(i, j + 1, k),
                   phi0.at(i, j - 1, k), phi0.at(i, j, k + 1), phi0.at(i, j, k - 1)};
  bool near = false;
#pragma unroll
  for (int n = 0; n < 6; n++) near = near || (nb[n] > z && c <= z) || (nb[n] <= z && c > z);
  if (!near) {
    dist0.at(i, j, k) = flag;
  } else {
    T delta = Tol<T>::zero;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T pp = nb[2 * d], pm = nb[2 * d + 1];
      const T d1 = fabs(pp - pm) / T(2), d2 = fabs(pp - c), d3 = fabs(c - pm);
      delta = fmax(fmax(fmax(d1, d2), d3), delta);
    }
    dist0.at(i, j, k) = c / delta;
  }
  LSM_CELL_LOOP_END
}

// :586-754
template <typename T>
__global__ void __launch_bounds__(256)
    k_subcell_rhs(Field<T> rhs, Field<const T> phi, Field<const T> phi0, Field<const T> dist0,
                  Field<const T> xp, Field<const T> yp, Field<const T> zp, Field<const T> xm,
                  Field<const T> ym, Field<const T> zm, Box fb, T max_dx, T flag) {
  LSM_CELL_LOOP_BEGIN(fb)
  const T p0 = phi0.at(i, j, k);
  const T s0 = (fabs(p0) > Tol<T>::zero) ? copysign(T(1), p0) : T(0);
  const T d0 = dist0.at(i, j, k);
  const T cur = phi.at(i, j, k);
  T r;
  if (d0 == flag) {
    if (fabs(cur) > Tol<T>::zero) {
      const T n2 = godunov_star_sq<T>(cur, xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k),
                                      xm.at(i, j, k), ym.at(i, j, k), zm.at(i, j, k));
      r = s0 - s0 * sqrt(n2);
    } else {
      r = T(0);
    }
  } else {
    r = d0 - s0 * fabs(cur) / max_dx;
  }
  rhs.at(i, j, k) = r;
  LSM_CELL_LOOP_END
}

// ============================================================================
// Boundary conditions  (lsm_boundary_conditions3d.f)
// ============================================================================
// One face: the threads cover the full ghost box in the two other directions
// (the reference loops there, :66-67), each writes the ghost cells of its line.
// mode 0 linear (:46-170), 1 signed linear (:193-330), 2 copy (:352-463)
template <typename T>
__global__ void __launch_bounds__(256) k_extrapolate(Field<T> phi, Box gb, Box fb, int loc, int mode) {
#pragma clang fp contract(off)  // pb + slope * dist rounded twice, as the reference does
  const int d = loc / 2, hi = loc % 2;
  const int u = (d + 1) % 3, v = (d + 2) % 3;
  const int glo[3] = {gb.ilo, gb.jlo, gb.klo}, ghi[3] = {gb.ihi, gb.jhi, gb.khi};
  const int flo[3] = {fb.ilo, fb.jlo, fb.klo}, fhi[3] = {fb.ihi, fb.jhi, fb.khi};
  const int pu = glo[u] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int pv = glo[v] + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (pu > ghi[u] || pv > ghi[v]) return;
  const int bnd = hi ? fhi[d] : flo[d];
  const int in = hi ? -1 : 1;
  int idx[3];
  idx[u] = pu;
  idx[v] = pv;
  idx[d] = bnd;
  const T pb = phi.at(idx[0], idx[1], idx[2]);
  idx[d] = bnd + in;
  const T p1 = phi.at(idx[0], idx[1], idx[2]);
  T slope = T(0);
  if (mode == 0) slope = pb - p1;
  if (mode == 1) slope = copysign(T(1), pb) * fabs(pb - p1);
  const int lo = hi ? bnd + 1 : glo[d], up = hi ? ghi[d] : bnd - 1;
  for (int t = lo; t <= up; t++) {
    idx[d] = t;
    if (mode == 2) {
      phi.at(idx[0], idx[1], idx[2]) = pb;
    } else {
      const T dist = T(hi ? (t - bnd) : (bnd - t));
      phi.at(idx[0], idx[1], idx[2]) = pb + slope * dist;
    }
  }
}

// ============================================================================
// Reductions (FP64 accumulation; one partial per block, then one block)
// ============================================================================
enum RedOp { RED_MAX = 0, RED_SUM = 1 };

__device__ __forceinline__ double red2(int op, double a, double b) {
  return op == RED_MAX ? (a > b ? a : b) : a + b;
}

// Block-wide reduction of one double per thread (256 threads = 4 waves of 64).
__device__ __forceinline__ double block_reduce(int op, double v, double ident) {
  __shared__ double s_w[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = red2(op, v, __shfl_xor(v, o, 64));
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  if ((t & 63) == 0) s_w[t >> 6] = v;
  __syncthreads();
  double r = ident;
  if (t == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    for (int w = 0; w < nw; w++) r = red2(op, r, s_w[w]);
  }
  return r;  // valid in thread 0
}

// Per-cell terms.  Each returns false when the cell is excluded.
struct RedArgs {
  int kind;
  real_t ix, iy, iz, eps, ooe, dV, max_dx_sq, vn_const;
  int cv_sgn;
  int cv_zero;     // control-volume cells equal to zero count too (voxel counts)
  int one_double;  // smoothed delta's (1 + cos) in FP64 (lsm_utilities3d) or LSMLIB_REAL (lsm_geometry3d)
  int ci, cj, ck;  // always-included cell (max-norm initial value)
};
enum RedKind {
  RK_ADV_DT = 1,
  RK_NORMAL_DT = 2,
  RK_CONST_NORMAL_DT = 3,
  RK_MAX_DIFF = 4,
  RK_VOL_INT_LT = 5,
  RK_VOL_INT_GT = 6,
  RK_VOL_REG_LT = 7,
  RK_VOL_REG_GT = 8,
  RK_AVE_ABS_DIFF = 9,
  RK_VOXEL_GT = 10,
  RK_VOXEL_LT = 11,
  RK_SURFACE = 12,
  RK_SURFACE_DELTA = 13
};

// Smoothed delta function (1 + cos(pi phi / eps)) / (2 eps) of the surface
// integrals and areas: the cosine is taken in LSMLIB_REAL, the (1 + cos) sum
// is FP64 in lsm_utilities3d.f.in (1.d0 + cos) and LSMLIB_REAL in
// lsm_geometry3d_fort.f.in (1 + cos); the products are FP64 (0.5d0 * ...).
template <typename T>
__device__ __forceinline__ T smoothed_delta(T p, T ooe, int one_double) {
#pragma clang fp contract(off)
  const T c = cos(T(3.14159265358979323846) * p * ooe);
  if (one_double) return T(0.5 * double(ooe) * (1.0 + double(c)));
  return T(0.5 * double(ooe) * double(T(1) + c));
}

// |grad phi| with the squares and sums rounded in LSMLIB_REAL
template <typename T>
__device__ __forceinline__ T norm3(T x, T y, T z) {
#pragma clang fp contract(off)
  return T(sqrt(double(x * x + y * y + z * z)));
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_reduce(RedArgs a, Box ib, Field<const T> f0, Field<const T> f1, Field<const T> f2,
             Field<const T> f3, Field<const T> f4, Field<const T> f5, Field<const T> f6,
             Field<const T> cv, double* partials) {
  const int op = (a.kind >= RK_VOL_INT_LT) ? RED_SUM : RED_MAX;
  const double ident = (op == RED_MAX) ? -1.0e308 : 0.0;
  double acc = ident;
  const int i = ib.ilo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = ib.jlo + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i <= ib.ihi && j <= ib.jhi) {
    for (int k = ib.klo + (int)blockIdx.z; k <= ib.khi; k += (int)gridDim.z) {
      bool inc = true;
      if (cv.p) {
        const T c = cv.at(i, j, k);
        inc = (a.cv_sgn > 0) ? (c > T(0)) : (c < T(0));
        if (a.cv_zero && c == T(0)) inc = true;
        if (a.kind == RK_MAX_DIFF && i == a.ci && j == a.cj && k == a.ck) inc = true;
      }
      if (!inc) continue;
      T v = T(0);
      switch (a.kind) {
        case RK_ADV_DT:  // lsm_utilities3d.f.in:306-308
          v = fabs(f0.at(i, j, k)) * T(a.ix) + fabs(f1.at(i, j, k)) * T(a.iy) + fabs(f2.at(i, j, k)) * T(a.iz);
          break;
        case RK_NORMAL_DT:
          v = h_over_dx<T>(fabs(f6.at(i, j, k)), f0.at(i, j, k), f1.at(i, j, k), f2.at(i, j, k),
                           f3.at(i, j, k), f4.at(i, j, k), f5.at(i, j, k), T(a.max_dx_sq), T(a.ix),
                           T(a.iy), T(a.iz));
          break;
        case RK_CONST_NORMAL_DT:
          v = h_over_dx<T>(fabs(T(a.vn_const)), f0.at(i, j, k), f1.at(i, j, k), f2.at(i, j, k),
                           f3.at(i, j, k), f4.at(i, j, k), f5.at(i, j, k), T(a.max_dx_sq), T(a.ix),
                           T(a.iy), T(a.iz));
          break;
        case RK_MAX_DIFF:  // lsm_utilities3d.f.in:56-64
          v = fabs(f0.at(i, j, k) - f1.at(i, j, k));
          break;
        case RK_VOL_INT_LT: {  // lsm_utilities3d.f.in:628-645
          const T p = f1.at(i, j, k), e = T(a.eps), x = p / e;
          if (p < -e) v = f0.at(i, j, k) * T(a.dV);
          else if (p < e)
            v = T(0.5) * (T(1) - x - T(0.31830988618379) * sin(T(3.14159265358979323846) * x)) *
                f0.at(i, j, k) * T(a.dV);
          else continue;
          break;
        }
        case RK_VOL_INT_GT: {  // lsm_utilities3d.f.in:713-731
          const T p = f1.at(i, j, k), e = T(a.eps), x = p / e;
          if (p > e) v = f0.at(i, j, k) * T(a.dV);
          else if (p > -e)
            v = T(0.5) * (T(1) + x + T(0.31830988618379) * sin(T(3.14159265358979323846) * x)) *
                f0.at(i, j, k) * T(a.dV);
          else continue;
          break;
        }
        case RK_VOL_REG_LT: {  // lsm_geometry3d_fort.f.in:315-325
          const T p = f0.at(i, j, k), e = T(a.eps), x = p / e;
          if (p < -e) v = T(a.dV);
          else if (p < e)
            v = T(0.5) * (T(1) - x - T(0.31830988618379) * sin(T(3.14159265358979323846) * x)) * T(a.dV);
          else continue;
          break;
        }
        case RK_VOL_REG_GT: {  // lsm_geometry3d_fort.f.in:405-412
          const T p = f0.at(i, j, k), e = T(a.eps), x = p / e;
          if (p > e) v = T(a.dV);
          else if (p > -e)
            v = T(0.5) * (T(1) + x + T(0.31830988618379) * sin(T(3.14159265358979323846) * x)) * T(a.dV);
          else continue;
          break;
        }
        case RK_AVE_ABS_DIFF:  // lsm_utilities3d.f.in:128-137
          v = fabs(f0.at(i, j, k) - f1.at(i, j, k));
          break;
        case RK_VOXEL_GT:  // lsm_utilities3d.f.in:184-190, :1889-1919
          if (!(f0.at(i, j, k) > T(0))) continue;
          v = T(1);
          break;
        case RK_VOXEL_LT:  // lsm_utilities3d.f.in:233-239, :1971-2001
          if (!(f0.at(i, j, k) < T(0))) continue;
          v = T(1);
          break;
        case RK_SURFACE: {  // lsm_utilities3d.f.in:826-836, lsm_geometry3d_fort.f.in:511-523
#pragma clang fp contract(off)
          const T p = f1.at(i, j, k);
          if (!(fabs(p) < T(a.eps))) continue;
          const T d = smoothed_delta<T>(p, T(a.ooe), a.one_double);
          const T ng = norm3<T>(f2.at(i, j, k), f3.at(i, j, k), f4.at(i, j, k));
          v = f0.p ? d * ng * f0.at(i, j, k) * T(a.dV) : d * ng * T(a.dV);
          break;
        }
        case RK_SURFACE_DELTA: {  // lsm_geometry3d_fort.f.in:597-604
#pragma clang fp contract(off)
          const T d = f0.at(i, j, k);
          if (!(d > T(0))) continue;
          v = d * norm3<T>(f2.at(i, j, k), f3.at(i, j, k), f4.at(i, j, k)) * T(a.dV);
          break;
        }
      }
      acc = red2(op, acc, (double)v);
    }
  }
  const double r = block_reduce(op, acc, ident);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = r;
}

__global__ void __launch_bounds__(256) k_reduce_final(const double* partials, int64_t n, int op, double* out) {
  const double ident = (op == RED_MAX) ? -1.0e308 : 0.0;
  double acc = ident;
  for (int64_t q = threadIdx.x; q < n; q += blockDim.x) acc = red2(op, acc, partials[q]);
  const double r = block_reduce(op, acc, ident);
  if (threadIdx.x == 0) out[0] = r;
}

}  // namespace lsm

// ============================================================================
// Host side: launch helpers
// ============================================================================
namespace {

using namespace lsm;
typedef real_t R;

template <typename K, typename... Args>
void launch_cells(Call& c, const char* kname, const Box& b, K kernel, Args... args) {
  if (!c.ok() || b.cells() == 0) return;
  hipLaunchKernelGGL(kernel, cell_grid(b), dim3(kBX, kBY, 1), 0, c.stream(), args...);
  c.launched(kname);
}

Box box6(const int* b) { return Box{b[0], b[1], b[2], b[3], b[4], b[5]}; }

Field<const R> no_field() {
  Field<const R> f;
  f.p = nullptr;
  f.ilo = f.jlo = f.klo = 0;
  f.sy = f.sz = 0;
  return f;
}

// Runs one reduction; returns the combined value (FP64).
double run_reduce(Call& c, RedArgs a, const Box& ib, Field<const R> f0, Field<const R> f1,
                  Field<const R> f2, Field<const R> f3, Field<const R> f4, Field<const R> f5,
                  Field<const R> f6, Field<const R> cv) {
  const int op = (a.kind >= RK_VOL_INT_LT) ? RED_SUM : RED_MAX;
  double ident = (op == RED_MAX) ? -1.0e308 : 0.0;
  if (!c.ok() || ib.cells() == 0) return ident;
  dim3 g = cell_grid(ib);
  if (g.z > 64) g.z = 64;
  const int64_t nb = (int64_t)g.x * g.y * g.z;
  double* part = reinterpret_cast<double*>(device_scratch((size_t)(nb + 1) * sizeof(double)));
  if (!c.ok()) return ident;
  hipLaunchKernelGGL(k_reduce<R>, g, dim3(kBX, kBY, 1), 0, c.stream(), a, ib, f0, f1, f2, f3, f4, f5, f6,
                     cv, part + 1);
  c.launched("k_reduce");
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream(), part + 1, nb, op, part);
  c.launched("k_reduce_final");
  double out = ident;
  check_hip(hipMemcpyAsync(&out, part, sizeof(double), hipMemcpyDeviceToHost, c.stream()),
            "hipMemcpyAsync(reduction result)");
  check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  return out;
}

}  // namespace

// ============================================================================
// Entry points (extern "C", reference prototypes; see include/lsm3d_amd.h)
// ============================================================================
#define BOXP(p)                                                                                \
  const int *ilo_##p, const int *ihi_##p, const int *jlo_##p, const int *jhi_##p, const int *klo_##p, \
      const int *khi_##p
#define BOX(p) lsm::box_of(ilo_##p, ihi_##p, jlo_##p, jhi_##p, klo_##p, khi_##p)

namespace {

template <int S>
void hj_pm_entry(const char* fn, R* xp, R* yp, R* zp, Box gp, R* xm, R* ym, R* zm, Box gm, const R* phi,
                 Box gphi, Box fb, R dx, R dy, R dz) {
  Call c(fn);
  constexpr int Rr = SchemeRadius<S>::R;
  require_inside(fn, "grad_phi_plus", gp, fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", gm, fb, 0, 0, 0);
  require_inside(fn, "phi", gphi, fb, Rr, Rr, Rr);
  auto fxp = c.out(xp, gp), fyp = c.out(yp, gp), fzp = c.out(zp, gp);
  auto fxm = c.out(xm, gm), fym = c.out(ym, gm), fzm = c.out(zm, gm);
  auto fphi = c.in(phi, gphi);
  launch_cells(c, "k_hj_pm", fb, k_hj_pm<R, S>, fxp, fyp, fzp, fxm, fym, fzm, fphi, fb, R(1) / dx,
               R(1) / dy, R(1) / dz);
  c.finish();
}

template <int S>
void hj_upwind_entry(const char* fn, R* gx, R* gy, R* gz, Box gg, const R* phi, Box gphi, const R* vx,
                     const R* vy, const R* vz, Box gv, Box fb, R dx, R dy, R dz) {
  Call c(fn);
  constexpr int Rr = SchemeRadius<S>::R;
  require_inside(fn, "grad_phi", gg, fb, 0, 0, 0);
  require_inside(fn, "phi", gphi, fb, Rr, Rr, Rr);
  require_inside(fn, "vel", gv, fb, 0, 0, 0);
  auto fx = c.out(gx, gg), fy = c.out(gy, gg), fz = c.out(gz, gg);
  auto fphi = c.in(phi, gphi);
  auto fvx = c.in(vx, gv), fvy = c.in(vy, gv), fvz = c.in(vz, gv);
  launch_cells(c, "k_hj_upwind", fb, k_hj_upwind<R, S>, fx, fy, fz, fphi, fvx, fvy, fvz, fb, R(1) / dx,
               R(1) / dy, R(1) / dz);
  c.finish();
}

void rk_entry(const char* fn, int form, R* out, Box go, const R* us, Box gs, const R* uc, Box gc,
              const R* rhs, Box gr, Box fb, R dt) {
  Call c(fn);
  require_inside(fn, "u_out", go, fb, 0, 0, 0);
  if (us) require_inside(fn, "u_stage", gs, fb, 0, 0, 0);
  require_inside(fn, "u_cur", gc, fb, 0, 0, 0);
  require_inside(fn, "rhs", gr, fb, 0, 0, 0);
  auto fo = c.out(out, go);
  Field<const R> fs = us ? c.in(us, gs) : no_field();
  auto fc = c.in(uc, gc);
  auto fr = c.in(rhs, gr);
  launch_cells(c, "k_rk", fb, k_rk<R>, form, fo, fs, fc, fr, fb, dt);
  c.finish();
}

void bc_entry(const char* fn, R* phi, Box gb, Box fb, int loc, int mode) {
  Call c(fn);
  if (loc < 0 || loc > 5) {  // the reference sets no ghost cells (:34-36)
    c.finish();
    return;
  }
  const int d = loc / 2;
  Box need = fb;  // the two cells next to the face must be interior cells
  if (d == 0 && fb.nx() < 2) raise_error(std::string(fn) + ": fill box needs 2 cells in x");
  if (d == 1 && fb.ny() < 2) raise_error(std::string(fn) + ": fill box needs 2 cells in y");
  if (d == 2 && fb.nz() < 2) raise_error(std::string(fn) + ": fill box needs 2 cells in z");
  require_inside(fn, "phi", gb, need, 0, 0, 0);
  auto f = c.out(phi, gb);
  if (c.ok()) {
    const int u = (d + 1) % 3, v = (d + 2) % 3;
    const int n[3] = {gb.nx(), gb.ny(), gb.nz()};
    dim3 g((unsigned)((n[u] + 63) / 64), (unsigned)((n[v] + 3) / 4), 1);
    hipLaunchKernelGGL(k_extrapolate<R>, g, dim3(64, 4, 1), 0, c.stream(), f, gb, fb, loc, mode);
    c.launched("k_extrapolate");
  }
  c.finish();
}

R max_of3(R a, R b, R c) {
  R m = a;
  if (b > m) m = b;
  if (c > m) m = c;
  return m;
}

}  // namespace

namespace lsm {
// ---- LSM3D_COMPUTE_STABLE_ADVECTION_DT without allocation -------------------
// The per-workgroup partial maxima live in a file-scope device array instead
// of the library's grown scratch: kNoAllocPartials doubles after one slot for
// the maximum, 512 KiB + 256 B, held by every GPU the library runs on until
// the process ends.  Its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int64_t kNoAllocPartials = 65536;
static __device__ __align__(256) double scratch[kNoAllocPartials + 32];

namespace noalloc {
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[64] = {nullptr};

double* noalloc_scratch(const char* fn) {
  int dev = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (error_pending()) return nullptr;
  if (dev < 0 || dev >= 64) {
    raise_error(std::string(fn) + ": device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      raise_error(std::string(fn) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
      return nullptr;
    }
    g_noalloc_ptr[dev] = static_cast<double*>(p);
  }
  return g_noalloc_ptr[dev];
}
}  // namespace

// dt = cfl / (max(-1, m) + tiny) from the reduced maximum m (lsm_utilities3d.f.in:297, 323),
// with m absent for an empty box
template <typename T>
__global__ void k_adv_dt_from_max(const double* m, int have, T cfl, T* dt) {
  T mx = T(-1.0);
  if (have && m[0] > (double)mx) mx = T(m[0]);
  dt[0] = cfl / (mx + Tol<T>::tiny);
}

}  // namespace lsm

extern "C" {

// ---- spatial derivatives (lsm_spatial_derivatives3d.h) -----------------------
void lsm3dhjeno1_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                  R* D1, BOXP(d1), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  (void)D1;
  hj_pm_entry<ENO1>("LSM3D_HJ_ENO1", xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), BOX(fb),
                    *dx, *dy, *dz);
}
void lsm3dhjeno2_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                  R* D1, BOXP(d1), R* D2, BOXP(d2), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  (void)D1;
  (void)D2;
  hj_pm_entry<ENO2>("LSM3D_HJ_ENO2", xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), BOX(fb),
                    *dx, *dy, *dz);
}
void lsm3dhjeno3_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                  R* D1, BOXP(d1), R* D2, BOXP(d2), R* D3, BOXP(d3), BOXP(fb), const R* dx, const R* dy,
                  const R* dz) {
  (void)D1;
  (void)D2;
  (void)D3;
  hj_pm_entry<ENO3>("LSM3D_HJ_ENO3", xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), BOX(fb),
                    *dx, *dy, *dz);
}
void lsm3dhjweno5_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                   R* D1, BOXP(d1), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  (void)D1;
  hj_pm_entry<WENO5>("LSM3D_HJ_WENO5", xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), BOX(fb),
                     *dx, *dy, *dz);
}
void lsm3dupwindhjeno1_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* vx, const R* vy,
                        const R* vz, BOXP(v), R* D1, BOXP(d1), BOXP(fb), const R* dx, const R* dy,
                        const R* dz) {
  (void)D1;
  hj_upwind_entry<ENO1>("LSM3D_UPWIND_HJ_ENO1", gx, gy, gz, BOX(g), phi, BOX(phi), vx, vy, vz, BOX(v),
                        BOX(fb), *dx, *dy, *dz);
}
void lsm3dupwindhjeno2_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* vx, const R* vy,
                        const R* vz, BOXP(v), R* D1, BOXP(d1), R* D2, BOXP(d2), BOXP(fb), const R* dx,
                        const R* dy, const R* dz) {
  (void)D1;
  (void)D2;
  hj_upwind_entry<ENO2>("LSM3D_UPWIND_HJ_ENO2", gx, gy, gz, BOX(g), phi, BOX(phi), vx, vy, vz, BOX(v),
                        BOX(fb), *dx, *dy, *dz);
}
void lsm3dupwindhjeno3_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* vx, const R* vy,
                        const R* vz, BOXP(v), R* D1, BOXP(d1), R* D2, BOXP(d2), R* D3, BOXP(d3),
                        BOXP(fb), const R* dx, const R* dy, const R* dz) {
  (void)D1;
  (void)D2;
  (void)D3;
  hj_upwind_entry<ENO3>("LSM3D_UPWIND_HJ_ENO3", gx, gy, gz, BOX(g), phi, BOX(phi), vx, vy, vz, BOX(v),
                        BOX(fb), *dx, *dy, *dz);
}
void lsm3dupwindhjweno5_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* vx,
                         const R* vy, const R* vz, BOXP(v), R* D1, BOXP(d1), BOXP(fb), const R* dx,
                         const R* dy, const R* dz) {
  (void)D1;
  hj_upwind_entry<WENO5>("LSM3D_UPWIND_HJ_WENO5", gx, gy, gz, BOX(g), phi, BOX(phi), vx, vy, vz, BOX(v),
                         BOX(fb), *dx, *dy, *dz);
}

// ---- central differences (lsm_spatial_derivatives3d.h:673-960) ------------------
// The factors are formed in double and rounded to LSMLIB_REAL, as the
// reference's d0 literals make them (:3058-3060, :3138-3140, :3220-3222).
static void central_grad_entry(const char* fn, int order, R* gx, R* gy, R* gz, Box gg, const R* phi, Box gphi,
                               Box fb, R dx, R dy, R dz) {
  Call c(fn);
  const int r = order / 2;
  require_inside(fn, "grad_phi", gg, fb, 0, 0, 0);
  require_inside(fn, "phi", gphi, fb, r, r, r);
  auto fx = c.out(gx, gg), fy = c.out(gy, gg), fz = c.out(gz, gg);
  auto fp = c.in(phi, gphi);
  if (order == 2)
    launch_cells(c, "k_central_grad2", fb, k_central_grad2<R>, fx, fy, fz, fp, fb, R(0.5 / (double)dx),
                 R(0.5 / (double)dy), R(0.5 / (double)dz));
  else
    launch_cells(c, "k_central_grad4", fb, k_central_grad4<R>, fx, fy, fz, fp, fb,
                 R(0.0833333333333333333333 / (double)dx), R(0.0833333333333333333333 / (double)dy),
                 R(0.0833333333333333333333 / (double)dz));
  c.finish();
}

void lsm3dcentralgradorder2_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), BOXP(fb), const R* dx,
                             const R* dy, const R* dz) {
  central_grad_entry("LSM3D_CENTRAL_GRAD_ORDER2", 2, gx, gy, gz, BOX(g), phi, BOX(phi), BOX(fb), *dx, *dy, *dz);
}

void lsm3dcentralgradorder4_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), BOXP(fb), const R* dx,
                             const R* dy, const R* dz) {
  central_grad_entry("LSM3D_CENTRAL_GRAD_ORDER4", 4, gx, gy, gz, BOX(g), phi, BOX(phi), BOX(fb), *dx, *dy, *dz);
}

void lsm3dlaplacianorder2_(R* lap, BOXP(l), const R* phi, BOXP(phi), BOXP(fb), const R* dx, const R* dy,
                           const R* dz) {
  const char* fn = "LSM3D_LAPLACIAN_ORDER2";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "laplacian_phi", BOX(l), fb, 0, 0, 0);
  require_inside(fn, "phi", BOX(phi), fb, 1, 1, 1);
  auto fl = c.out(lap, BOX(l));
  auto fp = c.in(phi, BOX(phi));
  const double x = *dx, y = *dy, z = *dz;
  launch_cells(c, "k_laplacian2", fb, k_laplacian2<R>, fl, fp, fb, R(1.0 / x / x), R(1.0 / y / y), R(1.0 / z / z));
  c.finish();
}

void lsm3dgradientmagnitude_(const R* gx, const R* gy, const R* gz, const R* mag_out, BOXP(g), BOXP(fb)) {
  R* mag = const_cast<R*>(mag_out);  // output, declared const by the reference
  const char* fn = "LSM3D_GRADIENT_MAGNITUDE";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "grad_phi", BOX(g), fb, 0, 0, 0);
  auto fm = c.out(mag, BOX(g));
  auto fx = c.in(gx, BOX(g)), fy = c.in(gy, BOX(g)), fz = c.in(gz, BOX(g));
  launch_cells(c, "k_grad_mag", fb, k_grad_mag<R>, fm, fx, fy, fz, fb);
  c.finish();
}

void lsm3dphiupwindgradf_(R* fx, R* fy, R* fz, BOXP(g), R* fxp, R* fyp, R* fzp, BOXP(gp), R* fxm, R* fym,
                          R* fzm, BOXP(gm), const R* phi, BOXP(phi), BOXP(fb)) {
  const char* fn = "LSM3D_PHI_UPWIND_GRAD_F";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "grad_F", BOX(g), fb, 0, 0, 0);
  require_inside(fn, "grad_F_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_F_minus", BOX(gm), fb, 0, 0, 0);
  require_inside(fn, "phi", BOX(phi), fb, 1, 1, 1);
  auto ox = c.out(fx, BOX(g)), oy = c.out(fy, BOX(g)), oz = c.out(fz, BOX(g));
  auto a = c.in(fxp, BOX(gp)), b = c.in(fyp, BOX(gp)), d = c.in(fzp, BOX(gp));
  auto e = c.in(fxm, BOX(gm)), f = c.in(fym, BOX(gm)), g = c.in(fzm, BOX(gm));
  auto p = c.in(phi, BOX(phi));
  launch_cells(c, "k_phi_upwind_grad", fb, k_phi_upwind_grad<R>, ox, oy, oz, a, b, d, e, f, g, p, fb);
  c.finish();
}

void lsm3daveragegradphi_(R* gx, R* gy, R* gz, BOXP(g), R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm,
                          BOXP(gm), BOXP(fb)) {
  const char* fn = "LSM3D_AVERAGE_GRAD_PHI";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "grad_phi", BOX(g), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  auto ox = c.out(gx, BOX(g)), oy = c.out(gy, BOX(g)), oz = c.out(gz, BOX(g));
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  launch_cells(c, "k_average_grad", fb, k_average_grad<R>, ox, oy, oz, a, b, d, e, f, g, fb);
  c.finish();
}

// ---- normals (lsm_geometry3d.h) and field extension (lsm_field_extension3d.h) ---
void lsm3dcomputeunitnormal_(R* nx, R* ny, R* nz, BOXP(n), const R* px, const R* py, const R* pz, BOXP(g),
                             BOXP(fb)) {
  const char* fn = "LSM3D_COMPUTE_UNIT_NORMAL";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "normal", BOX(n), fb, 0, 0, 0);
  require_inside(fn, "grad_phi", BOX(g), fb, 0, 0, 0);
  auto ox = c.out(nx, BOX(n)), oy = c.out(ny, BOX(n)), oz = c.out(nz, BOX(n));
  auto a = c.in(px, BOX(g)), b = c.in(py, BOX(g)), d = c.in(pz, BOX(g));
  launch_cells(c, "k_unit_normal", fb, k_unit_normal<R, false>, ox, oy, oz, a, b, d, no_field(), fb, R(0));
  c.finish();
}

void lsm3dcomputesignedunitnormal_(R* nx, R* ny, R* nz, BOXP(n), const R* px, const R* py, const R* pz, BOXP(g),
                                   const R* phi, BOXP(phi), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  const char* fn = "LSM3D_COMPUTE_SIGNED_UNIT_NORMAL";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "normal", BOX(n), fb, 0, 0, 0);
  require_inside(fn, "grad_phi", BOX(g), fb, 0, 0, 0);
  require_inside(fn, "phi", BOX(phi), fb, 0, 0, 0);
  auto ox = c.out(nx, BOX(n)), oy = c.out(ny, BOX(n)), oz = c.out(nz, BOX(n));
  auto a = c.in(px, BOX(g)), b = c.in(py, BOX(g)), d = c.in(pz, BOX(g));
  auto p = c.in(phi, BOX(phi));
  const R m = max_of3(*dx, *dy, *dz);
  launch_cells(c, "k_unit_normal", fb, k_unit_normal<R, true>, ox, oy, oz, a, b, d, p, fb, m * m);
  c.finish();
}

void lsm3dcomputefieldextensioneqnrhs_(R* rhs, BOXP(r), const R* S, BOXP(S), const R* phi, BOXP(phi),
                                       const R* sx, const R* sy, const R* sz, BOXP(gs), const R* nx, const R* ny,
                                       const R* nz, BOXP(n), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  const char* fn = "LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS";
  (void)S;
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "phi", BOX(phi), fb, 1, 1, 1);
  require_inside(fn, "grad_S_upwind", BOX(gs), fb, 0, 0, 0);
  require_inside(fn, "signed_normal", BOX(n), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto p = c.in(phi, BOX(phi));
  auto a = c.in(sx, BOX(gs)), b = c.in(sy, BOX(gs)), d = c.in(sz, BOX(gs));
  auto e = c.in(nx, BOX(n)), f = c.in(ny, BOX(n)), g = c.in(nz, BOX(n));
  // zero_level_set_cutoff = 3.0d0*max(dx,dy,dz) rounded to LSMLIB_REAL (:115)
  const R cutoff = R(3.0 * (double)max_of3(*dx, *dy, *dz));
  launch_cells(c, "k_field_ext_rhs", fb, k_field_ext_rhs<R>, fr, p, a, b, d, e, f, g, fb, cutoff);
  c.finish();
}

// Part 3: the velocity that turns the fused advection stage into the
// field-extension stage (see k_field_ext_velocity).
void lsm3d_amd_field_extension_velocity(R* vx, R* vy, R* vz, const int vel_gb[6], const R* nx, const R* ny,
                                        const R* nz, const int normal_gb[6], const R* phi, const int phi_gb[6],
                                        const int fbp[6], R dx, R dy, R dz) {
  const char* fn = "lsm3d_amd_field_extension_velocity";
  Call c(fn);
  const Box fb{fbp[0], fbp[1], fbp[2], fbp[3], fbp[4], fbp[5]};
  const Box gv{vel_gb[0], vel_gb[1], vel_gb[2], vel_gb[3], vel_gb[4], vel_gb[5]};
  const Box gn{normal_gb[0], normal_gb[1], normal_gb[2], normal_gb[3], normal_gb[4], normal_gb[5]};
  const Box gp{phi_gb[0], phi_gb[1], phi_gb[2], phi_gb[3], phi_gb[4], phi_gb[5]};
  require_inside(fn, "vel", gv, fb, 0, 0, 0);
  require_inside(fn, "signed_normal", gn, fb, 0, 0, 0);
  require_inside(fn, "phi", gp, fb, 1, 1, 1);
  auto ox = c.out(vx, gv), oy = c.out(vy, gv), oz = c.out(vz, gv);
  auto a = c.in(nx, gn), b = c.in(ny, gn), d = c.in(nz, gn);
  auto p = c.in(phi, gp);
  const R cutoff = R(3.0 * (double)max_of3(dx, dy, dz));
  launch_cells(c, "k_field_ext_velocity", fb, k_field_ext_velocity<R>, ox, oy, oz, a, b, d, p, fb, cutoff);
  c.finish();
}

// ---- level set equation RHS (lsm_level_set_evolution3d.h) --------------------
void lsm3dzerooutlevelseteqnrhs_(R* rhs, BOXP(r)) {
  const char* fn = "LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS";
  Call c(fn);
  Box gr = BOX(r);
  auto f = c.out(rhs, gr);
  launch_cells(c, "k_fill", gr, k_fill<R>, f, gr, R(0));
  c.finish();
}

void lsm3daddadvectiontermtolserhs_(R* rhs, BOXP(r), const R* px, const R* py, const R* pz, BOXP(g),
                                    const R* vx, const R* vy, const R* vz, BOXP(v), BOXP(fb)) {
  const char* fn = "LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "lse_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "grad_phi", BOX(g), fb, 0, 0, 0);
  require_inside(fn, "vel", BOX(v), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto fx = c.in(px, BOX(g)), fy = c.in(py, BOX(g)), fz = c.in(pz, BOX(g));
  auto ux = c.in(vx, BOX(v)), uy = c.in(vy, BOX(v)), uz = c.in(vz, BOX(v));
  launch_cells(c, "k_add_advection", fb, k_add_advection<R>, fr, fx, fy, fz, ux, uy, uz, fb);
  c.finish();
}

// Part 3: Lax-Friedrichs Hamiltonians (no reference counterpart).
void lsm3d_amd_add_lf_advection_term(R* rhs, const int rhs_gb[6], const R* xp, const R* yp, const R* zp,
                                     const int plus_gb[6], const R* xm, const R* ym, const R* zm,
                                     const int minus_gb[6], const R* vx, const R* vy, const R* vz,
                                     const int vel_gb[6], const int fbp[6], int alpha_mode, const R* alpha) {
  const char* fn = "lsm3d_amd_add_lf_advection_term";
  Call c(fn);
  const Box fb = box6(fbp);
  if (alpha_mode != 0 && alpha_mode != 1) raise_error(std::string(fn) + ": alpha_mode is 0 (local) or 1 (global)");
  if (alpha_mode == 1 && !alpha) raise_error(std::string(fn) + ": global alpha needs alpha[3]");
  if (alpha_mode == 1 && alpha && !(alpha[0] >= 0 && alpha[1] >= 0 && alpha[2] >= 0))
    raise_error(std::string(fn) + ": alpha must be non-negative");
  require_inside(fn, "lse_rhs", box6(rhs_gb), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", box6(plus_gb), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", box6(minus_gb), fb, 0, 0, 0);
  require_inside(fn, "vel", box6(vel_gb), fb, 0, 0, 0);
  if (!c.ok()) return;
  const R ax = alpha_mode ? alpha[0] : R(0), ay = alpha_mode ? alpha[1] : R(0), az = alpha_mode ? alpha[2] : R(0);
  auto fr = c.out(rhs, box6(rhs_gb));
  auto a = c.in(xp, box6(plus_gb)), b = c.in(yp, box6(plus_gb)), d = c.in(zp, box6(plus_gb));
  auto e = c.in(xm, box6(minus_gb)), f = c.in(ym, box6(minus_gb)), g = c.in(zm, box6(minus_gb));
  auto u = c.in(vx, box6(vel_gb)), v = c.in(vy, box6(vel_gb)), w = c.in(vz, box6(vel_gb));
  launch_cells(c, "k_add_lf_advection", fb, k_add_lf_advection<R>, fr, a, b, d, e, f, g, u, v, w, fb, alpha_mode,
               ax, ay, az);
  c.finish();
}

void lsm3d_amd_add_lf_normal_vel_term(R* rhs, const int rhs_gb[6], const R* xp, const R* yp, const R* zp,
                                      const int plus_gb[6], const R* xm, const R* ym, const R* zm,
                                      const int minus_gb[6], const R* vel_n, const int vel_n_gb[6], R vel_n_const,
                                      const int fbp[6], int alpha_mode, R alpha) {
  const char* fn = "lsm3d_amd_add_lf_normal_vel_term";
  Call c(fn);
  const Box fb = box6(fbp);
  if (alpha_mode != 0 && alpha_mode != 1) raise_error(std::string(fn) + ": alpha_mode is 0 (local) or 1 (global)");
  if (alpha_mode == 1 && !(alpha >= 0)) raise_error(std::string(fn) + ": alpha must be non-negative");
  require_inside(fn, "lse_rhs", box6(rhs_gb), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", box6(plus_gb), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", box6(minus_gb), fb, 0, 0, 0);
  if (vel_n) require_inside(fn, "vel_n", box6(vel_n_gb), fb, 0, 0, 0);
  if (!c.ok()) return;
  auto fr = c.out(rhs, box6(rhs_gb));
  auto a = c.in(xp, box6(plus_gb)), b = c.in(yp, box6(plus_gb)), d = c.in(zp, box6(plus_gb));
  auto e = c.in(xm, box6(minus_gb)), f = c.in(ym, box6(minus_gb)), g = c.in(zm, box6(minus_gb));
  Field<const R> n = vel_n ? c.in(vel_n, box6(vel_n_gb)) : no_field();
  launch_cells(c, "k_add_lf_normal", fb, k_add_lf_normal<R>, fr, a, b, d, e, f, g, n, vel_n_const, fb, alpha_mode,
               alpha);
  c.finish();
}

void lsm3daddnormalveltermtolserhs_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                    const R* xm, const R* ym, const R* zm, BOXP(gm), const R* vn,
                                    BOXP(v), BOXP(fb)) {
  const char* fn = "LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "lse_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  require_inside(fn, "vel_n", BOX(v), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  auto n = c.in(vn, BOX(v));
  launch_cells(c, "k_add_normal_vel", fb, k_add_normal_vel<R>, fr, a, b, d, e, f, g, n, R(0), fb);
  c.finish();
}

void lsm3daddconstnormalveltermtolserhs_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp,
                                         BOXP(gp), const R* xm, const R* ym, const R* zm, BOXP(gm),
                                         const R* vn, BOXP(fb)) {
  const char* fn = "LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "lse_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  if (std::fabs(*vn) >= Tol<R>::zero) {  // :300 the whole loop is skipped otherwise
    auto fr = c.out(rhs, BOX(r));
    auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
    auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
    launch_cells(c, "k_add_normal_vel", fb, k_add_normal_vel<R>, fr, a, b, d, e, f, g, no_field(), *vn, fb);
  }
  c.finish();
}

void lsm3daddconstcurvtermtolserhs_(const R* rhs_io, BOXP(r), const R* px, const R* py, const R* pz, BOXP(g1),
                                    const R* pxx, const R* pxy, const R* pxz, const R* pyy, const R* pyz,
                                    const R* pzz, BOXP(g2), const R* b, BOXP(fb)) {
  R* rhs = const_cast<R*>(rhs_io);  // in/out, declared const by the reference
  const char* fn = "LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "lse_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "grad_phi", BOX(g1), fb, 0, 0, 0);
  require_inside(fn, "grad2_phi", BOX(g2), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto x = c.in(px, BOX(g1)), y = c.in(py, BOX(g1)), z = c.in(pz, BOX(g1));
  auto xx = c.in(pxx, BOX(g2)), xy = c.in(pxy, BOX(g2)), xz = c.in(pxz, BOX(g2));
  auto yy = c.in(pyy, BOX(g2)), yz = c.in(pyz, BOX(g2)), zz = c.in(pzz, BOX(g2));
  launch_cells(c, "k_add_const_curv", fb, k_add_const_curv<R>, fr, x, y, z, xx, xy, xz, yy, yz, zz, fb, *b);
  c.finish();
}

void lsm3daddexternalandnormalveltermtolserhs_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp,
                                               BOXP(gp), const R* xm, const R* ym, const R* zm,
                                               BOXP(gm), const R* vn, const R* vx, const R* vy,
                                               const R* vz, BOXP(v), BOXP(fb)) {
  // The reference's loops run k = klo_rhs_fb..khi_rhs_gb, j = jlo_rhs_fb..jhi_rhs_gb,
  // i = ilo_rhs_gb..ihi_rhs_gb (lsm_level_set_evolution3d.f.in:620-622); the same box
  // is used here so results are identical, and every array must cover it.
  const char* fn = "LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS";
  Call c(fn);
  Box gr = BOX(r), fb = BOX(fb);
  Box lb{gr.ilo, gr.ihi, fb.jlo, gr.jhi, fb.klo, gr.khi};
  require_inside(fn, "lse_rhs", gr, lb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), lb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), lb, 0, 0, 0);
  require_inside(fn, "vel", BOX(v), lb, 0, 0, 0);
  auto fr = c.out(rhs, gr);
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  auto n = c.in(vn, BOX(v)), ux = c.in(vx, BOX(v)), uy = c.in(vy, BOX(v)), uz = c.in(vz, BOX(v));
  launch_cells(c, "k_add_ext_normal_vel", lb, k_add_ext_normal_vel<R>, fr, a, b, d, e, f, g, n, ux, uy, uz,
               lb);
  c.finish();
}

// ---- TVD Runge-Kutta (lsm_tvd_runge_kutta3d.h) --------------------------------
void lsm3drk1step_(R* un, BOXP(un), const R* uc, BOXP(uc), const R* rhs, BOXP(r), BOXP(fb), const R* dt) {
  rk_entry("LSM3D_RK1_STEP", RK_EULER, un, BOX(un), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), BOX(fb), *dt);
}
void lsm3dtvdrk2stage1_(R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs, BOXP(r), BOXP(fb),
                        const R* dt) {
  rk_entry("LSM3D_TVD_RK2_STAGE1", RK_EULER, u1, BOX(u1), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), BOX(fb),
           *dt);
}
void lsm3dtvdrk2stage2_(R* un, BOXP(un), const R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs,
                        BOXP(r), BOXP(fb), const R* dt) {
  rk_entry("LSM3D_TVD_RK2_STAGE2", RK2_STAGE2, un, BOX(un), u1, BOX(u1), uc, BOX(uc), rhs, BOX(r), BOX(fb),
           *dt);
}
void lsm3dtvdrk3stage1_(R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs, BOXP(r), BOXP(fb),
                        const R* dt) {
  rk_entry("LSM3D_TVD_RK3_STAGE1", RK_EULER, u1, BOX(u1), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), BOX(fb),
           *dt);
}
void lsm3dtvdrk3stage2_(R* u2, BOXP(u2), const R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs,
                        BOXP(r), BOXP(fb), const R* dt) {
  rk_entry("LSM3D_TVD_RK3_STAGE2", RK3_STAGE2, u2, BOX(u2), u1, BOX(u1), uc, BOX(uc), rhs, BOX(r), BOX(fb),
           *dt);
}
void lsm3dtvdrk3stage3_(R* un, BOXP(un), const R* u2, BOXP(u2), const R* uc, BOXP(uc), const R* rhs,
                        BOXP(r), BOXP(fb), const R* dt) {
  rk_entry("LSM3D_TVD_RK3_STAGE3", RK3_STAGE3, un, BOX(un), u2, BOX(u2), uc, BOX(uc), rhs, BOX(r), BOX(fb),
           *dt);
}

// ---- reinitialization (lsm_reinitialization3d.h) ----------------------------------
void lsm3dcomputereinitializationeqnrhs_(R* rhs, BOXP(r), const R* phi, BOXP(phi), const R* phi0,
                                         BOXP(phi0), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                         const R* xm, const R* ym, const R* zm, BOXP(gm), BOXP(fb),
                                         const R* dx, const R* dy, const R* dz, const int* use_phi0) {
  const char* fn = "LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  const int u0 = (*use_phi0 == 1);
  require_inside(fn, "reinit_rhs", BOX(r), fb, 0, 0, 0);
  if (u0) require_inside(fn, "phi0", BOX(phi0), fb, 0, 0, 0);
  else require_inside(fn, "phi", BOX(phi), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  Field<const R> fphi = u0 ? no_field() : c.in(phi, BOX(phi));
  Field<const R> fphi0 = u0 ? c.in(phi0, BOX(phi0)) : no_field();
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  R dx_sq = max_of3(*dx, *dy, *dz);  // :140-141
  dx_sq = dx_sq * dx_sq;
  launch_cells(c, "k_reinit_rhs", fb, k_reinit_rhs<R>, fr, fphi, fphi0, a, b, d, e, f, g, fb, dx_sq, u0);
  c.finish();
}

void lsm3dcomputeorthogonalizationeqnrhs_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                          const R* xm, const R* ym, const R* zm, BOXP(gm), const R* psi,
                                          BOXP(psi), const R* sxp, const R* syp, const R* szp, BOXP(sp),
                                          const R* sxm, const R* sym, const R* szm, BOXP(sm), BOXP(fb),
                                          const R* dx, const R* dy, const R* dz) {
  const char* fn = "LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "ortho_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  require_inside(fn, "psi", BOX(psi), fb, 0, 0, 0);
  require_inside(fn, "grad_psi_plus", BOX(sp), fb, 0, 0, 0);
  require_inside(fn, "grad_psi_minus", BOX(sm), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  auto fpsi = c.in(psi, BOX(psi));
  auto s1 = c.in(sxp, BOX(sp)), s2 = c.in(syp, BOX(sp)), s3 = c.in(szp, BOX(sp));
  auto s4 = c.in(sxm, BOX(sm)), s5 = c.in(sym, BOX(sm)), s6 = c.in(szm, BOX(sm));
  R dx_sq = max_of3(*dx, *dy, *dz);
  dx_sq = dx_sq * dx_sq;
  launch_cells(c, "k_ortho_rhs", fb, k_ortho_rhs<R>, fr, a, b, d, e, f, g, fpsi, s1, s2, s3, s4, s5, s6, fb,
               dx_sq);
  c.finish();
}

void lsm3dcomputedistanceforsubcellfix_(R* dist0, const R* phi0, BOXP(p0), BOXP(fb), const R* dx,
                                        const R* dy, const R* dz) {
  const char* fn = "LSM3D_COMPUTE_DISTANCE_FOR_SUBCELL_FIX";
  Call c(fn);
  Box fb = BOX(fb), g0 = BOX(p0);
  require_inside(fn, "phi0", g0, fb, 1, 1, 1);
  auto fd = c.out(dist0, g0);
  auto fp = c.in(phi0, g0);
  const R max_dx = max_of3(*dx, *dy, *dz);
  const R flag = R(-1000.0 * max_dx);
  launch_cells(c, "k_subcell_distance", fb, k_subcell_distance<R>, fd, fp, fb, flag);
  c.finish();
}

void lsm3dcomputereinitializationeqnrhssubcellfixorder1_(R* rhs, BOXP(r), const R* phi, BOXP(phi),
                                                         const R* phi0, const R* dist0, BOXP(p0),
                                                         const R* xp, const R* yp, const R* zp, BOXP(gp),
                                                         const R* xm, const R* ym, const R* zm, BOXP(gm),
                                                         BOXP(fb), const R* dx, const R* dy, const R* dz) {
  const char* fn = "LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_SUBCELL_FIX_ORDER1";
  Call c(fn);
  Box fb = BOX(fb);
  require_inside(fn, "reinit_rhs", BOX(r), fb, 0, 0, 0);
  require_inside(fn, "phi", BOX(phi), fb, 0, 0, 0);
  require_inside(fn, "phi0", BOX(p0), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", BOX(gp), fb, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", BOX(gm), fb, 0, 0, 0);
  auto fr = c.out(rhs, BOX(r));
  auto fphi = c.in(phi, BOX(phi));
  auto fp0 = c.in(phi0, BOX(p0)), fd0 = c.in(dist0, BOX(p0));
  auto a = c.in(xp, BOX(gp)), b = c.in(yp, BOX(gp)), d = c.in(zp, BOX(gp));
  auto e = c.in(xm, BOX(gm)), f = c.in(ym, BOX(gm)), g = c.in(zm, BOX(gm));
  const R max_dx = max_of3(*dx, *dy, *dz);
  const R flag = R(-1000.0 * max_dx);
  launch_cells(c, "k_subcell_rhs", fb, k_subcell_rhs<R>, fr, fphi, fp0, fd0, a, b, d, e, f, g, fb, max_dx, flag);
  c.finish();
}

// ---- reductions (lsm_utilities3d.h, lsm_geometry3d.h) ----------------------------
void lsm3dmaxnormdiff_(R* out, const R* f1, BOXP(f1), const R* f2, BOXP(f2), BOXP(ib)) {
  const char* fn = "LSM3D_MAX_NORM_DIFF";
  Call c(fn);
  Box ib = BOX(ib);
  require_inside(fn, "field1", BOX(f1), ib, 0, 0, 0);
  require_inside(fn, "field2", BOX(f2), ib, 0, 0, 0);
  auto a = c.in(f1, BOX(f1)), b = c.in(f2, BOX(f2));
  RedArgs ra{};
  ra.kind = RK_MAX_DIFF;
  double m = run_reduce(c, ra, ib, a, b, no_field(), no_field(), no_field(), no_field(), no_field(), no_field());
  c.finish();
  if (!lsm::error_pending()) *out = (ib.cells() > 0) ? R(m) : R(0);
}

void lsm3dmaxnormdiffcontrolvolume_(R* out, const R* f1, BOXP(f1), const R* f2, BOXP(f2), const R* cv,
                                    BOXP(cv), const int* cv_sgn, BOXP(ib)) {
  const char* fn = "LSM3D_MAX_NORM_DIFF_CONTROL_VOLUME";
  Call c(fn);
  Box ib = BOX(ib);
  require_inside(fn, "field1", BOX(f1), ib, 0, 0, 0);
  require_inside(fn, "field2", BOX(f2), ib, 0, 0, 0);
  require_inside(fn, "control_vol", BOX(cv), ib, 0, 0, 0);
  auto a = c.in(f1, BOX(f1)), b = c.in(f2, BOX(f2)), fc = c.in(cv, BOX(cv));
  RedArgs ra{};
  ra.kind = RK_MAX_DIFF;
  ra.cv_sgn = *cv_sgn;
  ra.ci = ib.ilo;
  ra.cj = ib.jlo;
  ra.ck = ib.klo;
  double m = run_reduce(c, ra, ib, a, b, no_field(), no_field(), no_field(), no_field(), no_field(), fc);
  c.finish();
  if (!lsm::error_pending()) *out = (ib.cells() > 0) ? R(m) : R(0);
}

static void adv_dt_entry(const char* fn, R* dt, const R* vx, const R* vy, const R* vz, Box gv, const R* cv,
                         Box gcv, int cv_sgn, Box ib, R dx, R dy, R dz, R cfl) {
  Call c(fn);
  require_inside(fn, "vel", gv, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  auto a = c.in(vx, gv), b = c.in(vy, gv), d = c.in(vz, gv);
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  RedArgs ra{};
  ra.kind = RK_ADV_DT;
  ra.ix = R(1) / dx;
  ra.iy = R(1) / dy;
  ra.iz = R(1) / dz;
  ra.cv_sgn = cv_sgn;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double m = run_reduce(c, ra, ib, a, b, d, no_field(), no_field(), no_field(), no_field(), fc);
  c.finish();
  R mx = R(-1.0);  // lsm_utilities3d.f.in:297
  if (m > (double)mx) mx = R(m);
  if (!lsm::error_pending()) *dt = cfl / (mx + Tol<R>::tiny);  // :323
}

void lsm3dcomputestableadvectiondt_(R* dt, const R* vx, const R* vy, const R* vz, BOXP(v), BOXP(ib),
                                    const R* dx, const R* dy, const R* dz, const R* cfl) {
  adv_dt_entry("LSM3D_COMPUTE_STABLE_ADVECTION_DT", dt, vx, vy, vz, BOX(v), nullptr, Box{}, 0, BOX(ib), *dx,
               *dy, *dz, *cfl);
}

void lsm3d_amd_compute_stable_advection_dt_noalloc(R* dt, const R* vx, const R* vy, const R* vz,
                                                   const int vel_gb[6], const int ibp[6], R dx, R dy, R dz, R cfl) {
  const char* fn = "lsm3d_amd_compute_stable_advection_dt_noalloc";
  Call c(fn);
  const Box ib = box6(ibp), gv = box6(vel_gb);
  if (!dt || !vx || !vy || !vz) raise_error(std::string(fn) + ": missing array");
  require_inside(fn, "vel", gv, ib, 0, 0, 0);
  if (!c.ok()) return;
  dim3 g = cell_grid(ib);
  const int64_t plane_blocks = (int64_t)g.x * g.y;
  if (ib.cells() > 0 && plane_blocks > kNoAllocPartials) {
    raise_error(std::string(fn) + ": box too large for the fixed scratch: ceil(nx/64) * ceil(ny/4) = " +
                std::to_string(plane_blocks) + " partial maxima per plane exceed " + std::to_string(kNoAllocPartials));
    return;
  }
  for (const void* q : {(const void*)dt, (const void*)vx, (const void*)vy, (const void*)vz})
    if (!is_device_pointer(q)) {
      raise_error(std::string(fn) + ": dt and the velocity arrays must be device memory");
      return;
    }
  double* sc = lsm::noalloc::noalloc_scratch(fn);
  if (!sc || !c.ok()) return;
  int64_t nb = 0;
  if (ib.cells() > 0) {
    int64_t gz = g.z;
    if (gz > 64) gz = 64;
    if (gz > kNoAllocPartials / plane_blocks) gz = kNoAllocPartials / plane_blocks;
    g.z = (unsigned)gz;
    nb = plane_blocks * gz;
    RedArgs ra{};
    ra.kind = RK_ADV_DT;
    ra.ix = R(1) / dx;
    ra.iy = R(1) / dy;
    ra.iz = R(1) / dz;
    ra.ci = ra.cj = ra.ck = INT32_MIN;
    Field<const R> a = make_field<const R>(vx, gv), b = make_field<const R>(vy, gv), d = make_field<const R>(vz, gv);
    hipLaunchKernelGGL(k_reduce<R>, g, dim3(kBX, kBY, 1), 0, c.stream(), ra, ib, a, b, d, no_field(), no_field(),
                       no_field(), no_field(), no_field(), sc + 1);
    c.launched("k_reduce");
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream(), sc + 1, nb, (int)RED_MAX, sc);
    c.launched("k_reduce_final");
  }
  hipLaunchKernelGGL(k_adv_dt_from_max<R>, dim3(1), dim3(1), 0, c.stream(), (const double*)sc, nb > 0 ? 1 : 0, cfl,
                     dt);
  c.launched("k_adv_dt_from_max");
  c.finish();
}

void lsm3dcomputestableadvectiondtcontrolvolume_(R* dt, const R* vx, const R* vy, const R* vz, BOXP(v),
                                                 const R* cv, BOXP(cv), const int* cv_sgn, BOXP(ib),
                                                 const R* dx, const R* dy, const R* dz, const R* cfl) {
  adv_dt_entry("LSM3D_COMPUTE_STABLE_ADVECTION_DT_CONTROL_VOLUME", dt, vx, vy, vz, BOX(v), cv, BOX(cv), *cv_sgn,
               BOX(ib), *dx, *dy, *dz, *cfl);
}

static void normal_dt_entry(const char* fn, R* dt, const R* vn, Box gv, R vn_const, const R* xp, const R* yp,
                            const R* zp, Box gp, const R* xm, const R* ym, const R* zm, Box gm, const R* cv,
                            Box gcv, int cv_sgn, Box ib, R dx, R dy, R dz, R cfl) {
  Call c(fn);
  if (vn) require_inside(fn, "vel_n", gv, ib, 0, 0, 0);
  require_inside(fn, "grad_phi_plus", gp, ib, 0, 0, 0);
  require_inside(fn, "grad_phi_minus", gm, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  Field<const R> fvn = vn ? c.in(vn, gv) : no_field();
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  auto a = c.in(xp, gp), b = c.in(yp, gp), d = c.in(zp, gp);
  auto e = c.in(xm, gm), f = c.in(ym, gm), g = c.in(zm, gm);
  RedArgs ra{};
  ra.kind = vn ? RK_NORMAL_DT : RK_CONST_NORMAL_DT;
  ra.ix = R(1) / dx;
  ra.iy = R(1) / dy;
  ra.iz = R(1) / dz;
  R mdx = max_of3(dx, dy, dz);
  ra.max_dx_sq = mdx * mdx;
  ra.vn_const = vn_const;
  ra.cv_sgn = cv_sgn;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double m = run_reduce(c, ra, ib, a, b, d, e, f, g, fvn, fc);
  c.finish();
  R mx = R(-1.0);
  if (m > (double)mx) mx = R(m);
  if (!lsm::error_pending()) *dt = cfl / (mx + Tol<R>::tiny);
}

void lsm3dcomputestablenormalveldt_(R* dt, const R* vn, BOXP(v), const R* xp, const R* yp, const R* zp,
                                    BOXP(gp), const R* xm, const R* ym, const R* zm, BOXP(gm), BOXP(ib),
                                    const R* dx, const R* dy, const R* dz, const R* cfl) {
  normal_dt_entry("LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT", dt, vn, BOX(v), R(0), xp, yp, zp, BOX(gp), xm, ym, zm,
                  BOX(gm), nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *cfl);
}

void lsm3dcomputestablenormalveldtcontrolvolume_(R* dt, const R* vn, BOXP(v), const R* xp, const R* yp,
                                                 const R* zp, BOXP(gp), const R* xm, const R* ym, const R* zm,
                                                 BOXP(gm), const R* cv, BOXP(cv), const int* cv_sgn, BOXP(ib),
                                                 const R* dx, const R* dy, const R* dz, const R* cfl) {
  normal_dt_entry("LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME", dt, vn, BOX(v), R(0), xp, yp, zp, BOX(gp),
                  xm, ym, zm, BOX(gm), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *cfl);
}

void lsm3dcomputestableconstnormalveldt_(R* dt, const R* vn, const R* xp, const R* yp, const R* zp, BOXP(gp),
                                         const R* xm, const R* ym, const R* zm, BOXP(gm), BOXP(ib),
                                         const R* dx, const R* dy, const R* dz, const R* cfl) {
  normal_dt_entry("LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT", dt, nullptr, Box{}, *vn, xp, yp, zp, BOX(gp), xm,
                  ym, zm, BOX(gm), nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *cfl);
}

void lsm3dcomputestableconstnormalveldtcontrolvolume_(R* dt, const R* vn, const R* xp, const R* yp, const R* zp,
                                                      BOXP(gp), const R* xm, const R* ym, const R* zm, BOXP(gm),
                                                      const R* cv, BOXP(cv), const int* cv_sgn, BOXP(ib),
                                                      const R* dx, const R* dy, const R* dz, const R* cfl) {
  normal_dt_entry("LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_CONTROL_VOLUME", dt, nullptr, Box{}, *vn, xp, yp, zp,
                  BOX(gp), xm, ym, zm, BOX(gm), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *cfl);
}

static void volume_entry(const char* fn, int kind, R* out, const R* F, Box gF, const R* phi, Box gphi,
                         const R* cv, Box gcv, int cv_sgn, Box ib, R dx, R dy, R dz, R eps) {
  Call c(fn);
  if (F) require_inside(fn, "F", gF, ib, 0, 0, 0);
  require_inside(fn, "phi", gphi, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  Field<const R> fF = F ? c.in(F, gF) : no_field();
  auto fp = c.in(phi, gphi);
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  RedArgs ra{};
  ra.kind = kind;
  ra.eps = eps;
  ra.dV = dx * dy * dz;  // dV = dx * dy * dz in LSMLIB_REAL
  ra.cv_sgn = cv_sgn;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double s;
  if (F) s = run_reduce(c, ra, ib, fF, fp, no_field(), no_field(), no_field(), no_field(), no_field(), fc);
  else s = run_reduce(c, ra, ib, fp, no_field(), no_field(), no_field(), no_field(), no_field(), no_field(), fc);
  c.finish();
  if (!lsm::error_pending()) *out = R(s);
}

void lsm3dvolumeintegralphilessthanzero_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi), BOXP(ib),
                                         const R* dx, const R* dy, const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO", RK_VOL_INT_LT, out, F, BOX(F), phi, BOX(phi),
               nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeintegralphigreaterthanzero_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi), BOXP(ib),
                                            const R* dx, const R* dy, const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO", RK_VOL_INT_GT, out, F, BOX(F), phi, BOX(phi),
               nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeregionphilessthanzero_(R* out, const R* phi, BOXP(phi), BOXP(ib), const R* dx, const R* dy,
                                       const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO", RK_VOL_REG_LT, out, nullptr, Box{}, phi, BOX(phi),
               nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeregionphigreaterthanzero_(R* out, const R* phi, BOXP(phi), BOXP(ib), const R* dx, const R* dy,
                                          const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO", RK_VOL_REG_GT, out, nullptr, Box{}, phi, BOX(phi),
               nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *eps);
}

void lsm3daveabsdiff_(R* out, const R* f1, BOXP(f1), const R* f2, BOXP(f2), BOXP(ib)) {
  const char* fn = "LSM3D_AVE_ABS_DIFF";
  Call c(fn);
  Box ib = BOX(ib);
  require_inside(fn, "field1", BOX(f1), ib, 0, 0, 0);
  require_inside(fn, "field2", BOX(f2), ib, 0, 0, 0);
  auto a = c.in(f1, BOX(f1)), b = c.in(f2, BOX(f2));
  RedArgs ra{};
  ra.kind = RK_AVE_ABS_DIFF;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double s = run_reduce(c, ra, ib, a, b, no_field(), no_field(), no_field(), no_field(), no_field(), no_field());
  c.finish();
  // lsm_utilities3d.f.in:141-145: the point count is the included box's
  const double n = (double)ib.cells();
  if (!lsm::error_pending()) *out = (n > 0) ? R(s / n) : R(0);
}

static void voxel_entry(const char* fn, int kind, int* count, const R* phi, Box gphi, const R* cv, Box gcv,
                        int cv_sgn, Box ib) {
  Call c(fn);
  require_inside(fn, "phi", gphi, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  auto fp = c.in(phi, gphi);
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  RedArgs ra{};
  ra.kind = kind;
  ra.cv_sgn = cv_sgn;
  ra.cv_zero = 1;  // lsm_utilities3d.f.in:1895-1896: control_vol .ge. 0 / .le. 0
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double s = run_reduce(c, ra, ib, fp, no_field(), no_field(), no_field(), no_field(), no_field(), no_field(), fc);
  c.finish();
  if (!lsm::error_pending()) *count = (int)s;
}

void lsm3dvoxelcountgreaterthanzero_(int* count, const R* phi, BOXP(g), BOXP(ib)) {
  voxel_entry("LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO", RK_VOXEL_GT, count, phi, BOX(g), nullptr, Box{}, 0,
              BOX(ib));
}
void lsm3dvoxelcountlessthanzero_(int* count, const R* phi, BOXP(g), BOXP(ib)) {
  voxel_entry("LSM3D_VOXEL_COUNT_LESS_THAN_ZERO", RK_VOXEL_LT, count, phi, BOX(g), nullptr, Box{}, 0, BOX(ib));
}
void lsm3dvoxelcountgreaterthanzerocontrolvolume_(int* count, const R* phi, BOXP(g), const R* cv, BOXP(cv),
                                                  const int* cv_sgn, BOXP(ib)) {
  voxel_entry("LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO_CONTROL_VOLUME", RK_VOXEL_GT, count, phi, BOX(g), cv, BOX(cv),
              *cv_sgn, BOX(ib));
}
void lsm3dvoxelcountlessthanzerocontrolvolume_(int* count, const R* phi, BOXP(g), const R* cv, BOXP(cv),
                                               const int* cv_sgn, BOXP(ib)) {
  voxel_entry("LSM3D_VOXEL_COUNT_LESS_THAN_ZERO_CONTROL_VOLUME", RK_VOXEL_LT, count, phi, BOX(g), cv, BOX(cv),
              *cv_sgn, BOX(ib));
}

// Surface integrals and areas: F (null for an area) times the smoothed delta
// function times |grad phi| over |phi| < eps; one_double selects the
// lsm_utilities3d form of the delta function (see smoothed_delta).
static void surface_entry(const char* fn, R* out, const R* F, Box gF, const R* phi, Box gphi, const R* px,
                          const R* py, const R* pz, Box gg, const R* cv, Box gcv, int cv_sgn, Box ib, R dx, R dy,
                          R dz, R eps, int one_double) {
  Call c(fn);
  if (F) require_inside(fn, "F", gF, ib, 0, 0, 0);
  require_inside(fn, "phi", gphi, ib, 0, 0, 0);
  require_inside(fn, "grad_phi", gg, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  Field<const R> fF = F ? c.in(F, gF) : no_field();
  auto fp = c.in(phi, gphi);
  auto a = c.in(px, gg), b = c.in(py, gg), d = c.in(pz, gg);
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  RedArgs ra{};
  ra.kind = RK_SURFACE;
  ra.eps = eps;
  ra.ooe = R(1.0 / (double)eps);  // one_over_epsilon = 1.d0/epsilon
  ra.dV = dx * dy * dz;
  ra.cv_sgn = cv_sgn;
  ra.one_double = one_double;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double s = run_reduce(c, ra, ib, fF, fp, a, b, d, no_field(), no_field(), fc);
  c.finish();
  if (!lsm::error_pending()) *out = R(s);
}

void lsm3dsurfaceintegral_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi), const R* px, const R* py,
                           const R* pz, BOXP(g), BOXP(ib), const R* dx, const R* dy, const R* dz, const R* eps) {
  surface_entry("LSM3D_SURFACE_INTEGRAL", out, F, BOX(F), phi, BOX(phi), px, py, pz, BOX(g), nullptr, Box{}, 0,
                BOX(ib), *dx, *dy, *dz, *eps, 1);
}
void lsm3dsurfaceintegralcontrolvolume_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi), const R* px,
                                        const R* py, const R* pz, BOXP(g), const R* cv, BOXP(cv),
                                        const int* cv_sgn, BOXP(ib), const R* dx, const R* dy, const R* dz,
                                        const R* eps) {
  surface_entry("LSM3D_SURFACE_INTEGRAL_CONTROL_VOLUME", out, F, BOX(F), phi, BOX(phi), px, py, pz, BOX(g), cv,
                BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps, 1);
}
void lsm3dsurfaceareazerolevelset_(R* out, const R* phi, BOXP(phi), const R* px, const R* py, const R* pz,
                                   BOXP(g), BOXP(ib), const R* dx, const R* dy, const R* dz, const R* eps) {
  surface_entry("LSM3D_SURFACE_AREA_ZERO_LEVEL_SET", out, nullptr, Box{}, phi, BOX(phi), px, py, pz, BOX(g),
                nullptr, Box{}, 0, BOX(ib), *dx, *dy, *dz, *eps, 0);
}
void lsm3dsurfaceareazerolevelsetcontrolvolume_(R* out, const R* phi, BOXP(phi), const R* px, const R* py,
                                                const R* pz, BOXP(g), const R* cv, BOXP(cv), const int* cv_sgn,
                                                BOXP(ib), const R* dx, const R* dy, const R* dz, const R* eps) {
  surface_entry("LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_CONTROL_VOLUME", out, nullptr, Box{}, phi, BOX(phi), px, py,
                pz, BOX(g), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps, 0);
}

static void surface_delta_entry(const char* fn, R* out, const R* dphi, Box gd, const R* px, const R* py,
                                const R* pz, Box gg, const R* cv, Box gcv, int cv_sgn, Box ib, R dx, R dy, R dz) {
  Call c(fn);
  require_inside(fn, "delta_phi", gd, ib, 0, 0, 0);
  require_inside(fn, "grad_phi", gg, ib, 0, 0, 0);
  if (cv) require_inside(fn, "control_vol", gcv, ib, 0, 0, 0);
  auto fd = c.in(dphi, gd);
  auto a = c.in(px, gg), b = c.in(py, gg), d = c.in(pz, gg);
  Field<const R> fc = cv ? c.in(cv, gcv) : no_field();
  RedArgs ra{};
  ra.kind = RK_SURFACE_DELTA;
  ra.dV = dx * dy * dz;
  ra.cv_sgn = cv_sgn;
  ra.ci = ra.cj = ra.ck = INT32_MIN;
  double s = run_reduce(c, ra, ib, fd, no_field(), a, b, d, no_field(), no_field(), fc);
  c.finish();
  if (!lsm::error_pending()) *out = R(s);
}

void lsm3dsurfaceareazerolevelsetdelta_(R* out, const R* dphi, BOXP(d), const R* px, const R* py, const R* pz,
                                        BOXP(g), BOXP(ib), const R* dx, const R* dy, const R* dz) {
  surface_delta_entry("LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA", out, dphi, BOX(d), px, py, pz, BOX(g), nullptr,
                      Box{}, 0, BOX(ib), *dx, *dy, *dz);
}
void lsm3dsurfaceareazerolevelsetdeltacontrolvolume_(R* out, const R* dphi, BOXP(d), const R* px, const R* py,
                                                     const R* pz, BOXP(g), const R* cv, BOXP(cv),
                                                     const int* cv_sgn, BOXP(ib), const R* dx, const R* dy,
                                                     const R* dz) {
  surface_delta_entry("LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA_CONTROL_VOLUME", out, dphi, BOX(d), px, py, pz,
                      BOX(g), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz);
}

void lsm3ddeltafunctionorder1_(const R* phi, const R* delta_out, BOXP(g), const R* nx, const R* ny, const R* nz,
                               BOXP(n), BOXP(fb), const R* dx, const R* dy, const R* dz) {
  R* delta = const_cast<R*>(delta_out);  // output, declared const by the reference
  const char* fn = "LSM3D_DELTA_FUNCTION_ORDER1";
  (void)nz;
  (void)dz;
  Call c(fn);
  Box fb = BOX(fb);
  // the reference reads phi at the six neighbours of every fill-box cell
  require_inside(fn, "phi", BOX(g), fb, 1, 1, 1);
  require_inside(fn, "norm_phi", BOX(n), fb, 0, 0, 0);
  auto fd = c.out(delta, BOX(g));
  auto fp = c.in(phi, BOX(g));
  auto fx = c.in(nx, BOX(n)), fy = c.in(ny, BOX(n));
  launch_cells(c, "k_delta_order1", fb, k_delta_order1<R>, fd, fp, fx, fy, fb, R(1.0 / (double)*dx),
               R(1.0 / (double)*dy));
  c.finish();
}

void lsm3dvolumeintegralphilessthanzerocontrolvolume_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi),
                                                      const R* cv, BOXP(cv), const int* cv_sgn, BOXP(ib),
                                                      const R* dx, const R* dy, const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO_CONTROL_VOLUME", RK_VOL_INT_LT, out, F, BOX(F), phi,
               BOX(phi), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeintegralphigreaterthanzerocontrolvolume_(R* out, const R* F, BOXP(F), const R* phi, BOXP(phi),
                                                         const R* cv, BOXP(cv), const int* cv_sgn, BOXP(ib),
                                                         const R* dx, const R* dy, const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME", RK_VOL_INT_GT, out, F, BOX(F), phi,
               BOX(phi), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeregionphilessthanzerocontrolvolume_(R* out, const R* phi, BOXP(phi), const R* cv, BOXP(cv),
                                                    const int* cv_sgn, BOXP(ib), const R* dx, const R* dy,
                                                    const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO_CONTROL_VOLUME", RK_VOL_REG_LT, out, nullptr, Box{}, phi,
               BOX(phi), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps);
}
void lsm3dvolumeregionphigreaterthanzerocontrolvolume_(R* out, const R* phi, BOXP(phi), const R* cv, BOXP(cv),
                                                       const int* cv_sgn, BOXP(ib), const R* dx, const R* dy,
                                                       const R* dz, const R* eps) {
  volume_entry("LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME", RK_VOL_REG_GT, out, nullptr, Box{},
               phi, BOX(phi), cv, BOX(cv), *cv_sgn, BOX(ib), *dx, *dy, *dz, *eps);
}

// ---- curvature (lsm_curvature3d.h) -------------------------------------------------
}  // extern "C"

namespace {

template <int ORDER, bool GAUSS>
void curvature_entry(const char* fn, R* kappa, Box gk, const R* phi, Box gphi, const R* px, const R* py,
                     const R* pz, const R* gm, Box gg, Box fb, R dx, R dy, R dz) {
  Call c(fn);
  require_inside(fn, "kappa", gk, fb, 0, 0, 0);
  require_inside(fn, "phi", gphi, fb, ORDER / 2, ORDER / 2, ORDER / 2);
  // order 4 differentiates phi_z along x, phi_x along y and phi_y along z
  require_inside(fn, "grad_phi", gg, fb, ORDER == 4 ? 2 : 0, ORDER == 4 ? 2 : 0, ORDER == 4 ? 2 : 0);
  auto fk = c.out(kappa, gk);
  auto fp = c.in(phi, gphi);
  auto fx = c.in(px, gg), fy = c.in(py, gg), fz = c.in(pz, gg), fg = c.in(gm, gg);
  launch_cells(c, "k_curvature", fb, k_curvature<R, ORDER, GAUSS>, fk, fp, fx, fy, fz, fg, fb,
               curv_coef<ORDER>(dx, dy, dz));
  c.finish();
}

}  // namespace

extern "C" {

void lsm3dcomputemeancurvatureorder2_(R* kappa, BOXP(k), const R* phi, BOXP(phi), const R* px, const R* py,
                                      const R* pz, const R* gm, BOXP(g), BOXP(fb), const R* dx, const R* dy,
                                      const R* dz) {
  curvature_entry<2, false>("LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2", kappa, BOX(k), phi, BOX(phi), px, py, pz, gm,
                            BOX(g), BOX(fb), *dx, *dy, *dz);
}
void lsm3dcomputegaussiancurvatureorder2_(R* kappa, BOXP(k), const R* phi, BOXP(phi), const R* px, const R* py,
                                          const R* pz, const R* gm, BOXP(g), BOXP(fb), const R* dx, const R* dy,
                                          const R* dz) {
  curvature_entry<2, true>("LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2", kappa, BOX(k), phi, BOX(phi), px, py, pz,
                           gm, BOX(g), BOX(fb), *dx, *dy, *dz);
}
void lsm3dcomputemeancurvatureorder4_(R* kappa, BOXP(k), const R* phi, BOXP(phi), const R* px, const R* py,
                                      const R* pz, const R* gm, BOXP(g), BOXP(fb), const R* dx, const R* dy,
                                      const R* dz) {
  curvature_entry<4, false>("LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4", kappa, BOX(k), phi, BOX(phi), px, py, pz, gm,
                            BOX(g), BOX(fb), *dx, *dy, *dz);
}
void lsm3dcomputegaussiancurvatureorder4_(R* kappa, BOXP(k), const R* phi, BOXP(phi), const R* px, const R* py,
                                          const R* pz, const R* gm, BOXP(g), BOXP(fb), const R* dx, const R* dy,
                                          const R* dz) {
  curvature_entry<4, true>("LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4", kappa, BOX(k), phi, BOX(phi), px, py, pz,
                           gm, BOX(g), BOX(fb), *dx, *dy, *dz);
}

void lsm3daddconstprecomputedcurvtermtolserhs_(const R* rhs_io, BOXP(r), R* kappa, BOXP(k), const R* gm,
                                               BOXP(g), const R* b, BOXP(fb)) {
  R* rhs = const_cast<R*>(rhs_io);  // in/out, declared const by the reference
  const char* fn = "LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS";
  Call c(fn);
  const Box gr = BOX(r), fb = BOX(fb);
  // lsm_level_set_evolution3d.f.in:514-516: k from the fill box's klo to the
  // ghost box's khi, j likewise, i over the whole ghost box
  const Box d{gr.ilo, gr.ihi, fb.jlo, gr.jhi, fb.klo, gr.khi};
  if (d.cells() > 0) {
    require_inside(fn, "lse_rhs", gr, d, 0, 0, 0);
    require_inside(fn, "kappa", BOX(k), d, 0, 0, 0);
    require_inside(fn, "grad_phi_mag", BOX(g), d, 0, 0, 0);
  }
  auto fr = c.out(rhs, gr);
  auto fk = c.in(kappa, BOX(k)), fg = c.in(gm, BOX(g));
  launch_cells(c, "k_precomputed_curv", d, k_precomputed_curv<R>, fr, fk, fg, d, *b);
  c.finish();
}

// ---- box copy (src/parallel/fortran/lsm_samrai_f77_utilities.h) -------------------
void lsm3dsamraiutilitiescopydata_(R* dst, BOXP(d), const R* src, BOXP(s), BOXP(fb)) {
  const char* fn = "LSM3D_SAMRAI_UTILITIES_COPY_DATA";
  Call c(fn);
  const Box fb = BOX(fb);
  require_inside(fn, "dst_data", BOX(d), fb, 0, 0, 0);
  require_inside(fn, "src_data", BOX(s), fb, 0, 0, 0);
  auto fd = c.out(dst, BOX(d));
  auto fs = c.in(src, BOX(s));
  launch_cells(c, "k_copy_box", fb, k_copy_box<R>, fd, fs, fb);
  c.finish();
}

// ---- boundary conditions (lsm_boundary_conditions3d.h) ---------------------------
void lsm3dlinearextrapolation_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_LINEAR_EXTRAPOLATION", phi, BOX(g), BOX(fb), *loc, 0);
}
void lsm3dsignedlinearextrapolation_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_SIGNED_LINEAR_EXTRAPOLATION", phi, BOX(g), BOX(fb), *loc, 1);
}
void lsm3dcopyextrapolation_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_COPY_EXTRAPOLATION", phi, BOX(g), BOX(fb), *loc, 2);
}
void lsm3dhomogeneousneumanneno1_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_HOMOGENEOUS_NEUMANN_ENO1", phi, BOX(g), BOX(fb), *loc, 2);
}
void lsm3dhomogeneousneumanneno2_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_HOMOGENEOUS_NEUMANN_ENO2", phi, BOX(g), BOX(fb), *loc, 2);
}
void lsm3dhomogeneousneumanneno3_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_HOMOGENEOUS_NEUMANN_ENO3", phi, BOX(g), BOX(fb), *loc, 2);
}
void lsm3dhomogeneousneumannweno5_(R* phi, BOXP(g), BOXP(fb), const int* loc) {
  bc_entry("LSM3D_HOMOGENEOUS_NEUMANN_WENO5", phi, BOX(g), BOX(fb), *loc, 2);
}

}  // extern "C"
