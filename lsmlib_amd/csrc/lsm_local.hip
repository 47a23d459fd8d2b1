// gfx950 kernels behind LSMLIB's narrow-band routines: the band builders of
// lsm_localization3d.f and the *LOCAL kernels, which visit the points named
// by index lists instead of a box (Peng et al.'s local level set method).
//
// Lists.  index_x/y/z(nlo:nhi) hold the (i, j, k) of each band point.  A
// routine given (nlo, nhi) visits entries nlo..nhi; entry l sits at element
// l - nlo of the arrays passed (the Fortran declares them
// index_x(nlo_index:nhi_index)).  Marks are the Fortran's integer*1, so they
// compare as signed bytes; a point takes part when narrow_band <= mark.
//
// Kernels.  One thread per list entry.  Each entry is checked against the
// boxes of every array it touches, with the stencil's reach; an entry outside
// is skipped and counted, and a call that completes synchronously reports
// skipped entries as an error (the reference reads and writes outside its
// arrays there).  Sums (the average gradient) accumulate in double precision.
//
// Band building.  The reference scans phi's box in Fortran order appending
// every point with |phi| < width, then, level by level, appends the unmarked
// neighbours of the previous level's points, in the order
// (i+1, i-1, j+1, j-1, k+1, k-1) of first discovery
// (lsm_localization3d.f:48-298).  Both orders are reproduced exactly:
//   level 0: ordered stream compaction -- per-block counts, one scan of the
//            block counts, then each block writes its points at its offset
//            (wave ballots give the order inside a block);
//   level l: every (parent, direction) pair that reaches an unmarked cell
//            bids its pair number into a per-cell key with atomicMin; the
//            pair holding the minimum is the first discovery; the winning
//            pairs are compacted in pair order as in level 0.
// The host reads the level sizes back (n_lo / n_hi are host outputs).
#include <climits>
#include <cmath>
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_local_step.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

using lsm::Box;
using lsm::Call;
using lsm::Field;

namespace lsm {

typedef int8_t nb_t;

__device__ __forceinline__ bool inside(const Box& b, int i, int j, int k) {
  return i >= b.ilo && i <= b.ihi && j >= b.jlo && j <= b.jhi && k >= b.klo && k <= b.khi;
}

// One launch's list entries.
struct Entries {
  const int* ix;
  const int* iy;
  const int* iz;
  int n;
  Field<const nb_t> nb;  // nb.p == nullptr: no mark test
  int mark;
  Box any;    // every entry: the arrays touched whatever the mark (narrow_band, a Dn output)
  Box valid;  // marked entries: the array boxes less the stencil's reach
  int* skipped;
};

__device__ __forceinline__ bool entry_ok(const Entries& e, int i, int j, int k, bool& marked) {
  if (!inside(e.any, i, j, k)) return false;
  marked = (e.nb.p == nullptr) || ((int)e.nb.at(i, j, k) <= e.mark);
  return !marked || inside(e.valid, i, j, k);
}

template <typename Op>
__global__ void __launch_bounds__(256) k_entries(Entries e, Op op) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= e.n) return;
  const int i = e.ix[t], j = e.iy[t], k = e.iz[t];
  bool marked = false;
  if (!entry_ok(e, i, j, k, marked)) {
    atomicAdd(e.skipped, 1);
    return;
  }
  op(i, j, k, marked);
}

// ---------------------------------------------------------------------------
// Per-entry operations.  `m`: the point passes the mark test.
// ---------------------------------------------------------------------------
template <typename T>
struct OpCopy {  // lsm_localization3d.f:724-776
  Field<T> dst;
  Field<const T> src;
  __device__ void operator()(int i, int j, int k, bool) const { dst.at(i, j, k) = src.at(i, j, k); }
};

template <typename T>
struct OpMaskMax {  // lsm_localization3d.f:660-718: max(mask, src)
  Field<T> dst;
  Field<const T> src, mask;
  __device__ void operator()(int i, int j, int k, bool) const {
    const T a = mask.at(i, j, k), b = src.at(i, j, k);
    dst.at(i, j, k) = (a > b) ? a : b;
  }
};

template <typename T>
struct OpZero {  // lsm_level_set_evolution3d_local.f.in:16-55
  Field<T> rhs;
  __device__ void operator()(int i, int j, int k, bool) const { rhs.at(i, j, k) = T(0); }
};

template <typename T>
struct OpCutoff {  // lsm_localization3d.f:482-562
  Field<T> rhs;
  Field<const T> phi;
  T beta, gamma, c1, c2;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T a = fabs(phi.at(i, j, k));
    T cut;
    if (a <= beta) {
      cut = T(1);
    } else if (a <= gamma) {
      const T t = a - gamma;
      cut = ((t * t) * (T(2) * a + c1)) / c2;
    } else {
      cut = T(0);
    }
    rhs.at(i, j, k) = cut * rhs.at(i, j, k);
  }
};

template <typename T>
struct OpRk {  // lsm_tvd_runge_kutta3d_local.f
  int form;
  Field<T> out;
  Field<const T> us, uc, rhs;
  T dt;
  __device__ void operator()(int i, int j, int k, bool m) const {
    if (!m) return;
    const T c = uc.at(i, j, k);
    out.at(i, j, k) = rk_combine<T>(form, c, us.p ? us.at(i, j, k) : c, rhs.at(i, j, k), dt);
  }
};

// lsm3dComputeDnLOCAL (lsm_spatial_derivatives3d_local.f.in:43-164):
// Dn = sgn (D(n-1)(c) - D(n-1)(c - off)) on marked points, 1e10 elsewhere
template <typename T>
struct OpDn {
  Field<T> dn;
  Field<const T> src;
  int oi, oj, ok;
  T sgn;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    dn.at(i, j, k) = m ? sgn * (src.at(i, j, k) - src.at(i - oi, j - oj, k - ok)) : T(1.e10);
  }
};

template <typename T>
struct OpEno1 {  // lsm_spatial_derivatives3d_local.f.in:290-300
  Field<T> plus, minus;
  Field<const T> d1;
  int di, dj, dk;
  T inv;
  __device__ void operator()(int i, int j, int k, bool m) const {
    if (!m) return;
    plus.at(i, j, k) = d1.at(i + di, j + dj, k + dk) * inv;
    minus.at(i, j, k) = d1.at(i, j, k) * inv;
  }
};

template <typename T>
struct OpEno2 {  // lsm_spatial_derivatives3d_local.f.in:571-601
  Field<T> plus, minus;
  Field<const T> d1, d2;
  int di, dj, dk;
  T inv;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T h = T(0.5);
    const T d2m = d2.at(i - di, j - dj, k - dk), d20 = d2.at(i, j, k), d2p = d2.at(i + di, j + dj, k + dk);
    const T d1p = d1.at(i + di, j + dj, k + dk), d10 = d1.at(i, j, k);
    plus.at(i, j, k) = (fabs(d20) < fabs(d2p)) ? (d1p - h * d20) * inv : (d1p - h * d2p) * inv;
    minus.at(i, j, k) = (fabs(d2m) < fabs(d20)) ? (d10 + h * d2m) * inv : (d10 + h * d20) * inv;
  }
};

template <typename T>
struct OpCentral2 {  // lsm_spatial_derivatives3d_local.f.in:826-840
  Field<T> gx, gy, gz;
  Field<const T> phi;
  T fx, fy, fz;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    gx.at(i, j, k) = (phi.at(i + 1, j, k) - phi.at(i - 1, j, k)) * fx;
    gy.at(i, j, k) = (phi.at(i, j + 1, k) - phi.at(i, j - 1, k)) * fy;
    gz.at(i, j, k) = (phi.at(i, j, k + 1) - phi.at(i, j, k - 1)) * fz;
  }
};

template <typename T>
struct OpCentral4 {  // lsm_spatial_derivatives3d_local.f.in:936-955
  Field<T> gx, gy, gz;
  Field<const T> phi;
  T fx, fy, fz;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T e = T(8);
    gx.at(i, j, k) = (((-phi.at(i + 2, j, k) + e * phi.at(i + 1, j, k)) + phi.at(i - 2, j, k)) -
                      e * phi.at(i - 1, j, k)) * fx;
    gy.at(i, j, k) = (((-phi.at(i, j + 2, k) + e * phi.at(i, j + 1, k)) + phi.at(i, j - 2, k)) -
                      e * phi.at(i, j - 1, k)) * fy;
    gz.at(i, j, k) = (((-phi.at(i, j, k + 2) + e * phi.at(i, j, k + 1)) + phi.at(i, j, k - 2)) -
                      e * phi.at(i, j, k - 1)) * fz;
  }
};

template <typename T, int ORDER, bool GAUSS>
struct OpCurvature {  // lsm_curvature3d_local.f.in
  Field<T> kappa;
  Field<const T> phi, px, py, pz, gm;
  CurvCoef<T> c;
  __device__ void operator()(int i, int j, int k, bool m) const {
    if (!m) return;
    kappa.at(i, j, k) = curvature_at<T, ORDER, GAUSS>(phi, px, py, pz, gm, i, j, k, c);
  }
};

template <typename T>
struct OpLaplacian {  // lsm_spatial_derivatives3d_local.f.in:1040-1055
  Field<T> lap;
  Field<const T> phi;
  T ix2, iy2, iz2;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T c = T(2) * phi.at(i, j, k);
    lap.at(i, j, k) = (ix2 * ((phi.at(i + 1, j, k) - c) + phi.at(i - 1, j, k)) +
                       iy2 * ((phi.at(i, j + 1, k) - c) + phi.at(i, j - 1, k))) +
                      iz2 * ((phi.at(i, j, k + 1) - c) + phi.at(i, j, k - 1));
  }
};

template <typename T>
struct OpGradMag {  // lsm_spatial_derivatives3d_local.f.in:1232-1258 (tolerance LSMLIB_ZERO_TOL)
  Field<T> mag;
  Field<const T> gx, gy, gz;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T tol = Tol<T>::zero, z = T(0);
    const T a = gx.at(i, j, k), b = gy.at(i, j, k), c = gz.at(i, j, k);
    T x2 = a * a, y2 = b * b, z2 = c * c;
    if (x2 < tol) x2 = z;
    if (y2 < tol) y2 = z;
    if (z2 < tol) z2 = z;
    const T t = sqrt((x2 + y2) + z2);
    mag.at(i, j, k) = (t < tol) ? z : t;
  }
};

template <typename T>
struct OpAdvection {  // lsm_level_set_evolution3d_local.f.in:130-142
  Field<T> rhs;
  Field<const T> px, py, pz, vx, vy, vz;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    rhs.at(i, j, k) = rhs.at(i, j, k) - ((vx.at(i, j, k) * px.at(i, j, k) + vy.at(i, j, k) * py.at(i, j, k)) +
                                         vz.at(i, j, k) * pz.at(i, j, k));
  }
};

template <typename T>
struct OpNormalVel {  // :236-272 (field vn) and :357-392 (constant vn: vnf.p == nullptr)
  Field<T> rhs;
  Field<const T> xp, yp, zp, xm, ym, zm, vnf;
  T vn_const;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T vn = vnf.p ? vnf.at(i, j, k) : vn_const;
    if (fabs(vn) >= Tol<T>::zero) {
      const T n2 = godunov_norm_sq<T>(vn, xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k), xm.at(i, j, k),
                                      ym.at(i, j, k), zm.at(i, j, k));
      rhs.at(i, j, k) = rhs.at(i, j, k) - vn * sqrt(n2);
    }
  }
};

template <typename T>
struct OpConstCurv {  // :490-518
  Field<T> rhs;
  Field<const T> px, py, pz, pxx, pxy, pxz, pyy, pyz, pzz;
  T b;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T curv = curvature_term<T>(px.at(i, j, k), py.at(i, j, k), pz.at(i, j, k), pxx.at(i, j, k),
                                     pxy.at(i, j, k), pxz.at(i, j, k), pyy.at(i, j, k), pyz.at(i, j, k),
                                     pzz.at(i, j, k), Tol<T>::zero);
    rhs.at(i, j, k) = rhs.at(i, j, k) + b * curv;
  }
};

template <typename T>
struct OpPrecomputedCurv {  // :593-601
  Field<T> rhs;
  Field<const T> kappa, gmag;
  T b;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    rhs.at(i, j, k) = rhs.at(i, j, k) + (b * kappa.at(i, j, k)) * gmag.at(i, j, k);
  }
};

template <typename T>
struct OpExtNormal {  // :712-800
  Field<T> rhs;
  Field<const T> xp, yp, zp, xm, ym, zm, vn, vx, vy, vz;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T vnc = vn.at(i, j, k), vxc = vx.at(i, j, k), vyc = vy.at(i, j, k), vzc = vz.at(i, j, k);
    const T px = ext_normal_select<T>(vxc, vnc, xp.at(i, j, k), xm.at(i, j, k));
    const T py = ext_normal_select<T>(vyc, vnc, yp.at(i, j, k), ym.at(i, j, k));
    const T pz = ext_normal_select<T>(vzc, vnc, zp.at(i, j, k), zm.at(i, j, k));
    const T n2 = (px * px + py * py) + pz * pz;
    T r = ((rhs.at(i, j, k) - vxc * px) - vyc * py) - vzc * pz;
    if (fabs(vnc) >= Tol<T>::zero && n2 >= Tol<T>::zero) r = r - vnc * sqrt(n2);
    rhs.at(i, j, k) = r;
  }
};

template <typename T>
struct OpReinit {  // lsm_reinitialization3d_local.f.in:50-320
  Field<T> rhs;
  Field<const T> phi, phi0, xp, yp, zp, xm, ym, zm;
  T dx_sq;
  int use_phi0;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T cur = use_phi0 ? phi0.at(i, j, k) : phi.at(i, j, k);
    T r = T(0);
    if (fabs(cur) >= Tol<T>::zero) {
      const T n2 = godunov_star_sq<T>(cur, xp.at(i, j, k), yp.at(i, j, k), zp.at(i, j, k), xm.at(i, j, k),
                                      ym.at(i, j, k), zm.at(i, j, k));
      const T sgn = use_phi0 ? cur / sqrt(cur * cur + dx_sq) : cur / sqrt(cur * cur + n2 * dx_sq);
      r = sgn * (T(1) - sqrt(n2));
    }
    rhs.at(i, j, k) = r;
  }
};

template <typename T>
struct OpOrtho {  // lsm_reinitialization3d_local.f.in:326-470
  Field<T> rhs;
  Field<const T> xp, yp, zp, xm, ym, zm, psi, sxp, syp, szp, sxm, sym, szm;
  T dx_sq;
  __device__ void operator()(int i, int j, int k, bool m) const {
#pragma clang fp contract(off)
    if (!m) return;
    const T g0 = T(0.5) * (sxp.at(i, j, k) + sxm.at(i, j, k));
    const T g1 = T(0.5) * (syp.at(i, j, k) + sym.at(i, j, k));
    const T g2 = T(0.5) * (szp.at(i, j, k) + szm.at(i, j, k));
    const T ng = sqrt((g0 * g0 + g1 * g1) + g2 * g2);
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
  }
};

// ---------------------------------------------------------------------------
// Reductions over list entries: a max or a sum, and the count of included
// entries; FP64 partials per block, then one block combines them.
// ---------------------------------------------------------------------------
template <typename T>
struct RedMaxDiff {  // lsm_utilities3d_local.f.in:16-92
  Field<const T> f1, f2;
  __device__ bool operator()(int i, int j, int k, bool m, double& v) const {
    if (!m) return false;
    v = (double)fabs(f1.at(i, j, k) - f2.at(i, j, k));
    return true;
  }
};

template <typename T>
struct RedAveGrad {  // lsm_spatial_derivatives3d_local.f.in:1140-1170
  Field<const T> phi;
  T fx, fy, fz;
  __device__ bool operator()(int i, int j, int k, bool m, double& v) const {
#pragma clang fp contract(off)
    if (!m) return false;
    const T x = (phi.at(i + 1, j, k) - phi.at(i - 1, j, k)) * fx;
    const T y = (phi.at(i, j + 1, k) - phi.at(i, j - 1, k)) * fy;
    const T z = (phi.at(i, j, k + 1) - phi.at(i, j, k - 1)) * fz;
    v = (double)sqrt((x * x + y * y) + z * z);
    return true;
  }
};

template <typename T>
struct RedAdvDt {  // lsm_utilities3d_local.f.in:262-275
  Field<const T> vx, vy, vz;
  T ix, iy, iz;
  __device__ bool operator()(int i, int j, int k, bool m, double& v) const {
#pragma clang fp contract(off)
    if (!m) return false;
    v = (double)((fabs(vx.at(i, j, k)) * ix + fabs(vy.at(i, j, k)) * iy) + fabs(vz.at(i, j, k)) * iz);
    return true;
  }
};

// Normal-velocity CFL.  Field vn (cv optional): lsm_utilities3d_local.f.in:
// 390-420 and :540-615, taken where the smoothed |grad phi| exceeds
// LSMLIB_ZERO_TOL.  Constant vn (vnf.p == nullptr): the box routine's H/dX,
// the formula the reference's :628-752 intends (it reads two variables it
// never assigns).
template <typename T>
struct RedNormalDt {
  Field<const T> xp, yp, zp, xm, ym, zm, vnf, cv;
  T avn_const, mdx2, ix, iy, iz;
  int cv_sgn;
  __device__ bool operator()(int i, int j, int k, bool m, double& v) const {
#pragma clang fp contract(off)
    if (!m) return false;
    if (cv.p) {
      const T c = cv.at(i, j, k);
      if (!((cv_sgn > 0) ? (c > T(0)) : (c < T(0)))) return false;
    }
    const T a = xp.at(i, j, k), b = yp.at(i, j, k), c = zp.at(i, j, k);
    const T d = xm.at(i, j, k), e = ym.at(i, j, k), f = zm.at(i, j, k);
    if (vnf.p) {
      const T px = fmax(fabs(a), fabs(d)), py = fmax(fabs(b), fabs(e)), pz = fmax(fabs(c), fabs(f));
      const T ng = sqrt(((px * px + py * py) + pz * pz) + mdx2);
      if (!(ng > Tol<T>::zero)) return false;
      v = (double)(fabs(vnf.at(i, j, k)) / ng * ((px * ix + py * iy) + pz * iz));
    } else {
      v = (double)h_over_dx<T>(avn_const, a, b, c, d, e, f, mdx2, ix, iy, iz);
    }
    return true;
  }
};

enum { RED_LIST_MAX = 0, RED_LIST_SUM = 1 };

__device__ __forceinline__ void block_reduce2(int op, double& v, double& c) {
  __shared__ double s_v[4], s_c[4];
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = (op == RED_LIST_MAX) ? fmax(v, w) : v + w;
    c += __shfl_xor(c, o, 64);
  }
  const int t = threadIdx.x;
  if ((t & 63) == 0) {
    s_v[t >> 6] = v;
    s_c[t >> 6] = c;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < (int)(blockDim.x / 64); w++) {
      v = (op == RED_LIST_MAX) ? fmax(v, s_v[w]) : v + s_v[w];
      c += s_c[w];
    }
  }
}

template <typename Op>
__global__ void __launch_bounds__(256) k_entries_reduce(Entries e, Op op, int red, double* partials) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double v = (red == RED_LIST_MAX) ? -1.0e308 : 0.0, c = 0.0;
  if (t < e.n) {
    const int i = e.ix[t], j = e.iy[t], k = e.iz[t];
    bool marked = false;
    if (!entry_ok(e, i, j, k, marked)) {
      atomicAdd(e.skipped, 1);
    } else {
      double x;
      if (op(i, j, k, marked, x)) {
        v = x;
        c = 1.0;
      }
    }
  }
  block_reduce2(red, v, c);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = v;
    partials[2 * blockIdx.x + 1] = c;
  }
}

__global__ void __launch_bounds__(256) k_entries_reduce_final(const double* partials, int nblocks, int red,
                                                              double* out) {
  double v = (red == RED_LIST_MAX) ? -1.0e308 : 0.0, c = 0.0;
  for (int q = threadIdx.x; q < nblocks; q += blockDim.x) {
    const double w = partials[2 * q];
    v = (red == RED_LIST_MAX) ? fmax(v, w) : v + w;
    c += partials[2 * q + 1];
  }
  block_reduce2(red, v, c);
  if (threadIdx.x == 0) {
    out[0] = v;
    out[1] = c;
  }
}

// ---------------------------------------------------------------------------
// Band building
// ---------------------------------------------------------------------------
constexpr int kChunkRounds = 32;
constexpr int kChunk = 256 * kChunkRounds;  // cells (or pairs) per block
constexpr int kBatch = 16;                  // rounds whose loads are issued together
constexpr int kPairRounds = 8;
constexpr int kPairChunk = 256 * kPairRounds;  // (parent, direction) pairs per block

enum { BAND_PHI = 0, BAND_MASK_NONNEG = 1, BAND_MASK_NONPOS = 2 };

// lsm_localization3d.f:232-255 and :428-455.  Bit 0: a band point; bit 1: an
// outer-layer point with phi <= 0; bit 2: an outer-layer point with phi > 0.
template <typename T>
__device__ __forceinline__ int band_bits(int mode, T v, T width, T width_inner) {
  if (mode == BAND_PHI) {
    const T a = fabs(v);
    const int in = (a < width) ? 1 : 0;
    const int outer = (in && a >= width_inner) ? 1 : 0;
    const int neg = (v <= T(0)) ? 1 : 0;
    return in | ((outer & neg) << 1) | ((outer & (neg ^ 1)) << 2);
  }
  if (mode == BAND_MASK_NONNEG) return (v >= T(0)) ? 1 : 0;
  return (v <= T(0)) ? 1 : 0;
}

// n / d for 0 <= n < 2^31 by a multiply-high and a shift, the invariant-
// divisor method of Granlund and Montgomery (PLDI 1994): s = ceil(log2 d),
// m = floor(2^32 (2^s - d) / d) + 1, n / d = (mulhi(m, n) + n) >> s (the sum
// stays below 2^32 because mulhi(m, n) <= n < 2^31).
struct FastDiv {
  uint32_t m, s;
  __host__ __device__ FastDiv() : m(1), s(0) {}
  __host__ explicit FastDiv(uint32_t d) {
    s = 0;
    while (s < 32 && ((uint64_t)1 << s) < d) s++;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << s) - d)) / d + 1);
  }
  __device__ __forceinline__ int div(int n) const {
    return (int)((__umulhi(m, (uint32_t)n) + (uint32_t)n) >> s);
  }
};

struct CellMap {  // linear cell index of a box, Fortran order, and back
  Box b;
  int nx, nxy;
  FastDiv dx, dxy;
  __device__ __forceinline__ void ijk(int c, int& i, int& j, int& k) const {
    const int kk = dxy.div(c);
    const int r = c - kk * nxy;
    const int jj = dx.div(r);
    i = b.ilo + (r - jj * nx);
    j = b.jlo + jj;
    k = b.klo + kk;
  }
  __device__ __forceinline__ int lin(int i, int j, int k) const {
    return (i - b.ilo) + nx * (j - b.jlo) + nxy * (k - b.klo);
  }
};

// Block-wide counts of up to three flags; the wave sums go through LDS.
__device__ __forceinline__ void block_sum3(int& a, int& b, int& c) {
  __shared__ int s[3][4];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  const int t = threadIdx.x;
  if ((t & 63) == 0) {
    s[0][t >> 6] = a;
    s[1][t >> 6] = b;
    s[2][t >> 6] = c;
  }
  __syncthreads();
  a = s[0][0] + s[0][1] + s[0][2] + s[0][3];
  b = s[1][0] + s[1][1] + s[1][2] + s[1][3];
  c = s[2][0] + s[2][1] + s[2][2] + s[2][3];
}

// Exclusive in-block ranks of the three flag bits of `bits` for one round of
// 256 items, in item order; `tot` returns the round's totals.  Ends with a
// barrier, so the LDS it uses can be reused by the next round.
__device__ __forceinline__ void block_rank3(int bits, int& r0, int& r1, int& r2, int& t0, int& t1, int& t2) {
  __shared__ int s[3][4];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long b0 = __ballot(bits & 1), b1 = __ballot(bits & 2), b2 = __ballot(bits & 4);
  if (lane == 0) {
    s[0][w] = __popcll(b0);
    s[1][w] = __popcll(b1);
    s[2][w] = __popcll(b2);
  }
  __syncthreads();
  int p0 = 0, p1 = 0, p2 = 0;
  for (int x = 0; x < w; x++) {
    p0 += s[0][x];
    p1 += s[1][x];
    p2 += s[2][x];
  }
  r0 = p0 + __popcll(b0 & lt);
  r1 = p1 + __popcll(b1 & lt);
  r2 = p2 + __popcll(b2 & lt);
  t0 = s[0][0] + s[0][1] + s[0][2] + s[0][3];
  t1 = s[1][0] + s[1][1] + s[1][2] + s[1][3];
  t2 = s[2][0] + s[2][1] + s[2][2] + s[2][3];
  __syncthreads();
}

// Level 0 scans phi's box in linear (Fortran) order; the field is indexed by
// the same linear cell number (its Field is made on that box).  Each thread
// issues the loads of kBatch rounds before using them.
template <typename T>
__global__ void __launch_bounds__(256)
    k_band_count(Field<const T> f, int ncell, int mode, T width, T width_inner, int* counts, nb_t* codes) {
  const int base = (int)blockIdx.x * kChunk;
  int a = 0, b = 0, c = 0;
  for (int r0 = 0; r0 < kChunkRounds; r0 += kBatch) {
    T v[kBatch];
#pragma unroll
    for (int r = 0; r < kBatch; r++) v[r] = f.p[min(base + (r0 + r) * 256 + (int)threadIdx.x, ncell - 1)];
#pragma unroll
    for (int r = 0; r < kBatch; r++) {
      const int x = base + (r0 + r) * 256 + (int)threadIdx.x;
      const int bits = (x < ncell) ? band_bits<T>(mode, v[r], width, width_inner) : 0;
      if (codes && x < ncell) codes[x] = (nb_t)bits;  // the cell's flags, read back by k_band_write_codes
      a += bits & 1;
      b += (bits >> 1) & 1;
      c += (bits >> 2) & 1;
    }
  }
  block_sum3(a, b, c);
  if (threadIdx.x == 0) {
    counts[3 * blockIdx.x] = a;
    counts[3 * blockIdx.x + 1] = b;
    counts[3 * blockIdx.x + 2] = c;
  }
}

// Exclusive scan, in place, of `nb` rows of three counts (one workgroup of
// 1024 threads: each thread sums a run of rows, the runs are scanned with wave
// shuffles and one LDS step over the 16 wave totals); totals[0..2] receive the
// sums.
__device__ __forceinline__ void scan3_body(int* counts, int nb, int* totals);

__global__ void __launch_bounds__(1024) k_scan3(int* counts, int nb, int* totals) { scan3_body(counts, nb, totals); }

// The same with the row count read from device memory.
__global__ void __launch_bounds__(1024) k_scan3_dev(int* counts, const int* nb, int* totals) {
  scan3_body(counts, *nb, totals);
}

__device__ __forceinline__ void scan3_body(int* counts, int nb, int* totals) {
  __shared__ int s_w[3][16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = (nb + 1023) / 1024;
  const int lo = min(nb, t * per), hi = min(nb, lo + per);
  int sum[3] = {0, 0, 0};
#pragma unroll 8
  for (int r = lo; r < hi; r++)
    for (int q = 0; q < 3; q++) sum[q] += counts[3 * r + q];
  int incl[3];
  for (int q = 0; q < 3; q++) {
    int v = sum[q];
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    incl[q] = v;
    if (lane == 63) s_w[q][w] = v;
  }
  __syncthreads();
  int run[3];
  for (int q = 0; q < 3; q++) {
    int before = 0;
    for (int x = 0; x < w; x++) before += s_w[q][x];
    run[q] = before + incl[q] - sum[q];
  }
  for (int r = lo; r < hi; r++)
    for (int q = 0; q < 3; q++) {
      const int v = counts[3 * r + q];
      counts[3 * r + q] = run[q];
      run[q] += v;
    }
  if (t == 1023)
    for (int q = 0; q < 3; q++) {
      int all = 0;
      for (int x = 0; x < 16; x++) all += s_w[q][x];
      totals[q] = all;
    }
}

// Writes level 0: narrow_band over phi's box (1 in the band, 0 elsewhere),
// the band points in Fortran order, and the outer-layer entries (minus side
// ascending from outer[0], plus side descending from outer[plus_top]).
// same_box: narrow_band's box is phi's, so it is indexed by the cell number too.
template <typename T>
__global__ void __launch_bounds__(256)
    k_band_write(Field<const T> f, CellMap cm, int ncell, int mode, T width, T width_inner, const int* offsets,
                 Field<nb_t> nb, int same_box, int* ix, int* iy, int* iz, int list_base, int* outer,
                 int plus_top) {
  const int base = (int)blockIdx.x * kChunk;
  int o0 = offsets[3 * blockIdx.x], o1 = offsets[3 * blockIdx.x + 1], o2 = offsets[3 * blockIdx.x + 2];
  for (int r0 = 0; r0 < kChunkRounds; r0 += kBatch) {
    T v[kBatch];
#pragma unroll
    for (int r = 0; r < kBatch; r++) v[r] = f.p[min(base + (r0 + r) * 256 + (int)threadIdx.x, ncell - 1)];
    for (int r = 0; r < kBatch; r++) {
      const int x = base + (r0 + r) * 256 + (int)threadIdx.x;
      int bits = 0, i = 0, j = 0, k = 0;
      if (x < ncell) {
        bits = band_bits<T>(mode, v[r], width, width_inner);
        const nb_t mk = (nb_t)(bits & 1);
        if (same_box) {
          nb.p[x] = mk;
        } else {
          cm.ijk(x, i, j, k);
          nb.at(i, j, k) = mk;
        }
      }
      int r0_, r1_, r2_, t0_, t1_, t2_;
      block_rank3(bits, r0_, r1_, r2_, t0_, t1_, t2_);
      if (bits & 1) {
        cm.ijk(x, i, j, k);
        const int p = o0 + r0_;
        ix[p] = i;
        iy[p] = j;
        iz[p] = k;
        if (bits & 2) outer[o1 + r1_] = list_base + p;
        if (bits & 4) outer[plus_top - (o2 + r2_)] = list_base + p;
      }
      o0 += t0_;
      o1 += t1_;
      o2 += t2_;
    }
  }
}

// Level 0 from the flags k_band_count stored in narrow_band (narrow_band on
// phi's box): each thread takes 32 consecutive cells of the workgroup's
// chunk, so one block scan per chunk ranks them; the band points are written
// in cell order as k_band_write writes them, and narrow_band is left 1 on the
// band and 0 elsewhere (outer-layer flags cleared).  VEC: narrow_band's
// address is 16-byte aligned (two 16-byte loads per thread).
template <bool VEC>
__global__ void __launch_bounds__(256)
    k_band_write_codes(nb_t* codes, CellMap cm, int ncell, const int* offsets, int* ix, int* iy, int* iz,
                       int list_base, int* outer, int plus_top, const int* err) {
  __shared__ int s_w[3][4];
  const bool write = (err == nullptr) || (*err == 0);
  const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
  const int c0 = (int)blockIdx.x * kChunk + t * 32;
  uint32_t band = 0, oneg = 0, opos = 0;
  if (VEC && c0 + 32 <= ncell) {
    const uint4* q = reinterpret_cast<const uint4*>(codes + c0);
    const uint4 u0 = q[0], u1 = q[1];
    const uint32_t wv[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
    for (int q8 = 0; q8 < 8; q8++)
#pragma unroll
      for (int bq = 0; bq < 4; bq++) {
        const uint32_t code = (wv[q8] >> (8 * bq)) & 0xffu;
        const int bit = q8 * 4 + bq;
        band |= (code & 1u) << bit;
        oneg |= ((code >> 1) & 1u) << bit;
        opos |= ((code >> 2) & 1u) << bit;
      }
  } else {
    for (int q = 0; q < 32; q++) {
      const int x = c0 + q;
      const uint32_t code = (x < ncell) ? (uint32_t)(uint8_t)codes[x] : 0u;
      band |= (code & 1u) << q;
      oneg |= ((code >> 1) & 1u) << q;
      opos |= ((code >> 2) & 1u) << q;
    }
  }
  // exclusive ranks of the three counts in thread (= cell) order
  const int n0 = __popc(band), n1 = __popc(oneg), n2 = __popc(opos);
  int i0 = n0, i1 = n1, i2 = n2;
  for (int o = 1; o < 64; o <<= 1) {
    const int a0 = __shfl_up(i0, o, 64), a1 = __shfl_up(i1, o, 64), a2 = __shfl_up(i2, o, 64);
    if (lane >= o) {
      i0 += a0;
      i1 += a1;
      i2 += a2;
    }
  }
  if (lane == 63) {
    s_w[0][w] = i0;
    s_w[1][w] = i1;
    s_w[2][w] = i2;
  }
  __syncthreads();
  int p0 = offsets[3 * blockIdx.x] + i0 - n0, p1 = offsets[3 * blockIdx.x + 1] + i1 - n1,
      p2 = offsets[3 * blockIdx.x + 2] + i2 - n2;
  for (int x = 0; x < w; x++) {
    p0 += s_w[0][x];
    p1 += s_w[1][x];
    p2 += s_w[2][x];
  }
  uint32_t m = band;
  while (m) {
    const int q = __ffs((int)m) - 1;
    m &= m - 1;
    const int x = c0 + q;
    int i, j, k;
    cm.ijk(x, i, j, k);
    if (write) {
      ix[p0] = i;
      iy[p0] = j;
      iz[p0] = k;
    }
    if ((oneg >> q) & 1u) {
      if (write) outer[p1] = list_base + p0;
      p1++;
      codes[x] = nb_t(1);
    }
    if ((opos >> q) & 1u) {
      if (write) outer[plus_top - p2] = list_base + p0;
      p2++;
      codes[x] = nb_t(1);
    }
    p0++;
  }
}

// Neighbour q (0..5: i+1, i-1, j+1, j-1, k+1, k-1) of (i, j, k) inside the
// narrow-band box, lsm_localization3d.f:96-155.
__device__ __forceinline__ bool neighbour(int q, int i, int j, int k, const Box& b, int& a, int& bb, int& c) {
  a = i;
  bb = j;
  c = k;
  switch (q) {
    case 0: if (i >= b.ihi) return false; a = i + 1; break;
    case 1: if (i <= b.ilo) return false; a = i - 1; break;
    case 2: if (j >= b.jhi) return false; bb = j + 1; break;
    case 3: if (j <= b.jlo) return false; bb = j - 1; break;
    case 4: if (k >= b.khi) return false; c = k + 1; break;
    default: if (k <= b.klo) return false; c = k - 1; break;
  }
  return true;
}

struct Parents {
  const int* ix;
  const int* iy;
  const int* iz;  // entry 0 = the level's first point
  int npairs;     // 6 per parent
};

__global__ void __launch_bounds__(256) k_nbr_bid(Parents P, Field<const nb_t> nb, CellMap cm, int* key) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= P.npairs) return;
  const int m = p / 6, q = p - 6 * m;
  int a, b, c;
  if (neighbour(q, P.ix[m], P.iy[m], P.iz[m], cm.b, a, b, c) && nb.at(a, b, c) == 0)
    atomicMin(&key[cm.lin(a, b, c)], p);
}

__global__ void __launch_bounds__(256)
    k_nbr_count(Parents P, Field<const nb_t> nb, CellMap cm, const int* key, unsigned char* win, int* counts) {
  const int base = (int)blockIdx.x * kPairChunk;
  int n = 0, z0 = 0, z1 = 0;
  for (int r = 0; r < kPairRounds; r++) {
    const int p = base + r * 256 + (int)threadIdx.x;
    if (p < P.npairs) {
      const int m = p / 6, q = p - 6 * m;
      int a, b, c;
      const bool w = neighbour(q, P.ix[m], P.iy[m], P.iz[m], cm.b, a, b, c) && nb.at(a, b, c) == 0 &&
                     key[cm.lin(a, b, c)] == p;
      win[p] = w ? 1 : 0;
      n += w;
    }
  }
  block_sum3(n, z0, z1);
  if (threadIdx.x == 0) {
    counts[3 * blockIdx.x] = n;
    counts[3 * blockIdx.x + 1] = 0;
    counts[3 * blockIdx.x + 2] = 0;
  }
}

// Appends the winners in pair order, marks them, and clears their keys.
// write == 0 only clears the keys (the list is full: the call fails).
__global__ void __launch_bounds__(256)
    k_nbr_write(Parents P, Field<nb_t> nb, CellMap cm, int* key, const unsigned char* win, const int* offsets,
                int* ox, int* oy, int* oz, nb_t mark, int write) {
  const int base = (int)blockIdx.x * kPairChunk;
  int o = offsets[3 * blockIdx.x];
  for (int r = 0; r < kPairRounds; r++) {
    const int p = base + r * 256 + (int)threadIdx.x;
    int bits = 0, a = 0, b = 0, c = 0;
    if (p < P.npairs && win[p]) {
      const int m = p / 6, q = p - 6 * m;
      neighbour(q, P.ix[m], P.iy[m], P.iz[m], cm.b, a, b, c);
      bits = 1;
    }
    int r0_, r1_, r2_, t0_, t1_, t2_;
    block_rank3(bits, r0_, r1_, r2_, t0_, t1_, t2_);
    if (bits) {
      key[cm.lin(a, b, c)] = INT_MAX;
      if (write) {
        const int pos = o + r0_;
        ox[pos] = a;
        oy[pos] = b;
        oz[pos] = c;
        nb.at(a, b, c) = mark;
      }
    }
    o += t0_;
  }
}

// ---------------------------------------------------------------------------
// Band building with the level sizes kept on the device (the narrow-band
// model's step): the same kernels' work on the same order of entries, with
// every size read from a BandState instead of the host, so a band build
// needs no host synchronisation.  Kernels sized by the state run on fixed
// grids and stride over their work.
// ---------------------------------------------------------------------------
struct BandPrev {  // the level limits a band build leaves alone when level 0 is empty
  int lo[8], hi[8];
};

struct BandState {
  int count;         // next free list entry (absolute index)
  int err;           // a list or level was too long
  int nbr;           // level 0 is not empty: the neighbour levels are built
  int npairs, nchunks, plo, fits;  // the level being built
  int tot[3];        // level 0: band points, outer minus, outer plus
  int n_lo[8], n_hi[8];
  double dt;         // the narrow-band model's step: its dt
};

__global__ void k_band_level0(BandState* bs, const int* totals, int nlo, int cap, int ocap, int levels,
                              BandPrev prev) {
  const int t0 = totals[0], t1 = totals[1], t2 = totals[2];
  bs->tot[0] = t0;
  bs->tot[1] = t1;
  bs->tot[2] = t2;
  bs->err = (t0 > cap || (int64_t)t1 + t2 > ocap) ? 1 : 0;
  const int count = nlo + (bs->err ? 0 : t0);
  bs->count = count;
  bs->n_lo[0] = nlo;
  bs->nbr = (count > nlo) && !bs->err;
  bs->n_hi[0] = bs->nbr ? count - 1 : count;
  for (int l = 1; l <= levels && l < 8; l++) {  // kept when no neighbour level is built
    bs->n_lo[l] = prev.lo[l];
    bs->n_hi[l] = prev.hi[l];
  }
}

__global__ void k_band_level_setup(BandState* bs, int l, int nlo) {
  int np = 0;
  if (bs->nbr && !bs->err) {
    const int plo = bs->n_lo[l - 1], phi = bs->n_hi[l - 1];
    if (plo >= nlo && phi < bs->count && phi >= plo) {
      const int64_t np64 = 6 * ((int64_t)phi - plo + 1);
      if (np64 > INT_MAX)
        bs->err = 1;
      else
        np = (int)np64;
    }
    bs->plo = plo;
  }
  bs->npairs = np;
  bs->nchunks = (np + kPairChunk - 1) / kPairChunk;
}

__global__ void __launch_bounds__(256) k_nbr_bid_dev(const BandState* bs, const int* ix, const int* iy,
                                                     const int* iz, int nlo, Field<const nb_t> nb, CellMap cm,
                                                     int* key) {
  const int np = bs->npairs, off = bs->plo - nlo;
  for (int p = (int)(blockIdx.x * blockDim.x + threadIdx.x); p < np; p += (int)(gridDim.x * blockDim.x)) {
    const int m = off + p / 6, q = p % 6;
    int a, b, c;
    if (neighbour(q, ix[m], iy[m], iz[m], cm.b, a, b, c) && nb.at(a, b, c) == 0) atomicMin(&key[cm.lin(a, b, c)], p);
  }
}

// The winners of a level's bids: the pair holding its cell's key.  The test
// is the same before and during k_nbr_write_dev (only a winner writes its own
// cell's mark and key).
__device__ __forceinline__ bool nbr_wins(int p, int off, const int* ix, const int* iy, const int* iz,
                                         const Field<const nb_t>& nb, const CellMap& cm, const int* key, int& a,
                                         int& b, int& c) {
  const int m = off + p / 6, q = p % 6;
  return neighbour(q, ix[m], iy[m], iz[m], cm.b, a, b, c) && nb.at(a, b, c) == 0 && key[cm.lin(a, b, c)] == p;
}

__global__ void __launch_bounds__(256) k_nbr_count_dev(const BandState* bs, const int* ix, const int* iy,
                                                       const int* iz, int nlo, Field<const nb_t> nb, CellMap cm,
                                                       const int* key, int* counts) {
  const int np = bs->npairs, nch = bs->nchunks, off = bs->plo - nlo;
  for (int ch = (int)blockIdx.x; ch < nch; ch += (int)gridDim.x) {  // uniform per workgroup
    int n = 0, z0 = 0, z1 = 0;
    for (int r = 0; r < kPairRounds; r++) {
      const int p = ch * kPairChunk + r * 256 + (int)threadIdx.x;
      int a, b, c;
      if (p < np && nbr_wins(p, off, ix, iy, iz, nb, cm, key, a, b, c)) n++;
    }
    block_sum3(n, z0, z1);
    if (threadIdx.x == 0) {
      counts[3 * ch] = n;
      counts[3 * ch + 1] = 0;
      counts[3 * ch + 2] = 0;
    }
    __syncthreads();  // block_sum3's LDS is read again by the next chunk
  }
}

__global__ void __launch_bounds__(256)
    k_nbr_write_dev(BandState* bs, const int* totals, const int* offsets, int* ix, int* iy, int* iz, int nlo,
                    int nhi, Field<nb_t> nbw, CellMap cm, int* key, nb_t mark) {
  const int np = bs->npairs, nch = bs->nchunks, off = bs->plo - nlo;
  const int base_entry = bs->count;  // absolute index of the level's first entry
  const bool fits = (int64_t)base_entry + totals[0] - 1 <= (int64_t)nhi;
  const Field<const nb_t> nb = make_field<const nb_t>(nbw.p, cm.b);
  for (int ch = (int)blockIdx.x; ch < nch; ch += (int)gridDim.x) {
    int o = offsets[3 * ch];
    for (int r = 0; r < kPairRounds; r++) {
      const int p = ch * kPairChunk + r * 256 + (int)threadIdx.x;
      int a = 0, b = 0, c = 0;
      const int bits = (p < np && nbr_wins(p, off, ix, iy, iz, nb, cm, key, a, b, c)) ? 1 : 0;
      int r0_, r1_, r2_, t0_, t1_, t2_;
      block_rank3(bits, r0_, r1_, r2_, t0_, t1_, t2_);
      if (bits) {
        if (fits) {
          const int pos = base_entry - nlo + o + r0_;
          ix[pos] = a;
          iy[pos] = b;
          iz[pos] = c;
          nbw.at(a, b, c) = mark;
        }
        key[cm.lin(a, b, c)] = INT_MAX;
      }
      o += t0_;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) bs->fits = fits ? 1 : 0;
}

__global__ void k_band_level_finish(BandState* bs, const int* totals, int l) {
  if (!bs->nbr) return;  // the level limits keep their previous values
  int added = 0;
  if (bs->npairs > 0) {
    if (!bs->fits) bs->err = 1;
    else added = totals[0];
  }
  bs->count += added;
  if (bs->count > bs->n_hi[l - 1] + 1) {
    bs->n_lo[l] = bs->n_hi[l - 1] + 1;
    bs->n_hi[l] = bs->count - 1;
  } else {
    bs->n_lo[l] = -1;
    bs->n_hi[l] = -1;
  }
}

__global__ void __launch_bounds__(256) k_fill_int(int* p, int64_t n, int v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// lsm_localization3d.f:303-378: the six fill-box face planes across the box
// of narrow_band.  blockIdx.z: face (0..5).
__global__ void __launch_bounds__(256) k_mark_layer(Field<nb_t> nb, Box gb, Box fb, nb_t mark) {
  const int face = blockIdx.z, d = face / 2;
  const int plane = (face & 1) ? (d == 0 ? fb.ihi : d == 1 ? fb.jhi : fb.khi)
                               : (d == 0 ? fb.ilo : d == 1 ? fb.jlo : fb.klo);
  const int u = (int)(blockIdx.x * 64 + threadIdx.x % 64), v = (int)(blockIdx.y * 4 + threadIdx.x / 64);
  int i, j, k;
  if (d == 0) { i = plane; j = gb.jlo + u; k = gb.klo + v; }
  else if (d == 1) { i = gb.ilo + u; j = plane; k = gb.klo + v; }
  else { i = gb.ilo + u; j = gb.jlo + v; k = plane; }
  if (inside(gb, i, j, k)) nb.at(i, j, k) = mark;
}

// lsm_localization3d.f:568-654: bit 0 / 1 = a point with phi > 0 / <= 0 seen
template <typename T>
__global__ void __launch_bounds__(256)
    k_check_outer(const int* outer, int n, const int* ix, const int* iy, const int* iz, int nlo, int nhi,
                  Field<const T> phi, Box gphi, int shift, int* bits, int* skipped) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int v = 0;  // 1: phi > 0 seen, 2: phi <= 0 seen
  if (t < n) {
    const int l = outer[t];
    if (l < nlo || l > nhi) {
      atomicAdd(skipped, 1);
    } else {
      const int i = ix[l - nlo], j = iy[l - nlo], k = iz[l - nlo];
      if (!inside(gphi, i, j, k))
        atomicAdd(skipped, 1);
      else
        v = phi.at(i, j, k) > T(0) ? 1 : 2;
    }
  }
  // one atomic per wave: the OR of the wave's signs
  const unsigned long long pos = __ballot(v & 1), neg = __ballot(v & 2);
  if ((threadIdx.x & 63) == 0 && (pos | neg)) atomicOr(bits, ((pos ? 1 : 0) | (neg ? 2 : 0)) << shift);
}

// ---------------------------------------------------------------------------
// The narrow-band curvature model's stages fused per level-0 point (the
// routine chain and why the values agree: lsm_local_step.h).
// ---------------------------------------------------------------------------
template <typename T>
struct LocalCurvArgs {
  const int* ix;  // the band lists, entry l at element l
  const int* iy;
  const int* iz;
  const BandState* bs;  // level 0: entries bs->n_lo[0] .. bs->n_hi[0]
  Field<const nb_t> nb;
  Field<const T> ps;  // stage input
  Field<const T> pc;  // u_cur (stage 2)
  Field<T> rhs;       // stage 1: lse_rhs
  Field<T> out;       // stage 2: phi_next; RK1: phi_stage1
  Box gb;             // the arrays' box
  Box valid;          // gb less two cells: every read of a marked point
  int mark_fb, mark_d1, mark_d2;
  int normal, curv;          // the a > 0 and b > 0 branches of the model
  T vn, avn, b;              // a, |a|, b in the library precision
  T inv[3];                  // 1/dx: ENO2 (lsm_spatial_derivatives3d_local.f.in:571-601)
  T half[3];                 // 0.5/dx: central differences (:826-840)
  T mdx2;                    // max(dx)^2: the CFL H/dx
  T beta, gamma, c1, c2;     // the cut-off (lsm_localization3d.f:482-562)
  const T* dt;               // the step's dt on the device
  double* partials;          // stage 1 with the normal term: per-block max of H/dx
  int* skipped;
};

// The stage's right-hand side at a point that passes the fill-box mark; h
// receives the CFL speed H/dx when the normal term is on.  Returns false when
// a value the routine chain reads would be left unwritten by it.
template <typename T>
__device__ __forceinline__ bool local_curv_rhs(const LocalCurvArgs<T>& a, int i, int j, int k, T& rhs, double& h) {
#pragma clang fp contract(off)  // every product and sum rounded as the reference's
  auto P = [&](int di, int dj, int dk) -> T { return a.ps.at(i + di, j + dj, k + dk); };
  auto NB = [&](int di, int dj, int dk) -> int { return (int)a.nb.at(i + di, j + dj, k + dk); };
  const T us = P(0, 0, 0);
  T r = T(0);  // LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL
  if (a.normal) {
    T pl[3], mi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
      // D1 at offsets -1..2 and D2 at -1..1 along the axis (lsm3dComputeDnLOCAL:
      // D1(c) = phi(c) - phi(c - e), D2(c) = -(D1(c) - D1(c + e)), 1e10 where
      // narrow_band exceeds the D1 / D2 mark)
      T d1[4], d2[3];
#pragma unroll
      for (int q = -1; q <= 2; q++)
        d1[q + 1] = (NB(q * ex, q * ey, q * ez) <= a.mark_d1)
                        ? P(q * ex, q * ey, q * ez) - P((q - 1) * ex, (q - 1) * ey, (q - 1) * ez)
                        : T(1.e10);
#pragma unroll
      for (int q = -1; q <= 1; q++)
        d2[q + 1] = (NB(q * ex, q * ey, q * ez) <= a.mark_d2) ? -(d1[q + 1] - d1[q + 2]) : T(1.e10);
      const T hf = T(0.5), iv = a.inv[d];
      pl[d] = (fabs(d2[1]) < fabs(d2[2])) ? (d1[2] - hf * d2[1]) * iv : (d1[2] - hf * d2[2]) * iv;
      mi[d] = (fabs(d2[0]) < fabs(d2[1])) ? (d1[1] + hf * d2[0]) * iv : (d1[1] + hf * d2[1]) * iv;
    }
    if (fabs(a.vn) >= Tol<T>::zero)  // LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL
      r = r - a.vn * sqrt(godunov_norm_sq<T>(a.vn, pl[0], pl[1], pl[2], mi[0], mi[1], mi[2]));
    h = (double)h_over_dx<T>(a.avn, pl[0], pl[1], pl[2], mi[0], mi[1], mi[2], a.mdx2, a.inv[0], a.inv[1], a.inv[2]);
  }
  if (a.curv) {
    // first derivatives on levels 0-1 (D1 mark) at the point and its six
    // neighbours, second derivatives on level 0 (D2 mark) at the point
    if (NB(0, 0, 0) > a.mark_d2 || NB(1, 0, 0) > a.mark_d1 || NB(-1, 0, 0) > a.mark_d1 ||
        NB(0, 1, 0) > a.mark_d1 || NB(0, -1, 0) > a.mark_d1 || NB(0, 0, 1) > a.mark_d1 || NB(0, 0, -1) > a.mark_d1)
      return false;
    const T hx = a.half[0], hy = a.half[1], hz = a.half[2];
    const T gx = (P(1, 0, 0) - P(-1, 0, 0)) * hx;
    const T gy = (P(0, 1, 0) - P(0, -1, 0)) * hy;
    const T gz = (P(0, 0, 1) - P(0, 0, -1)) * hz;
    const T gx_ip = (P(2, 0, 0) - us) * hx, gx_im = (us - P(-2, 0, 0)) * hx;
    const T gy_jp = (P(0, 2, 0) - us) * hy, gy_jm = (us - P(0, -2, 0)) * hy;
    const T gz_kp = (P(0, 0, 2) - us) * hz, gz_km = (us - P(0, 0, -2)) * hz;
    const T gy_ip = (P(1, 1, 0) - P(1, -1, 0)) * hy, gy_im = (P(-1, 1, 0) - P(-1, -1, 0)) * hy;
    const T gz_ip = (P(1, 0, 1) - P(1, 0, -1)) * hz, gz_im = (P(-1, 0, 1) - P(-1, 0, -1)) * hz;
    const T gz_jp = (P(0, 1, 1) - P(0, 1, -1)) * hz, gz_jm = (P(0, -1, 1) - P(0, -1, -1)) * hz;
    // the last writer of each second derivative: xx, xy, xz = d/dx of phi_x,
    // phi_y, phi_z; yy, yz = d/dy of phi_y, phi_z; zz = d/dz of phi_z
    const T xx = (gx_ip - gx_im) * hx, xy = (gy_ip - gy_im) * hx, xz = (gz_ip - gz_im) * hx;
    const T yy = (gy_jp - gy_jm) * hy, yz = (gz_jp - gz_jm) * hy, zz = (gz_kp - gz_km) * hz;
    r = r + a.b * curvature_term<T>(gx, gy, gz, xx, xy, xz, yy, yz, zz, Tol<T>::zero);
  }
  const T av = fabs(us);  // LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL
  T cut;
  if (av <= a.beta) {
    cut = T(1);
  } else if (av <= a.gamma) {
    const T t = av - a.gamma;
    cut = ((t * t) * (T(2) * av + a.c1)) / a.c2;
  } else {
    cut = T(0);
  }
  rhs = cut * r;
  return true;
}

template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_local_curv_stage(LocalCurvArgs<T> a) {
  const int lo = a.bs->n_lo[0], n = a.bs->n_hi[0] - lo + 1;
  double h = -1.0e308;
  for (int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); t < n; t += (int)(gridDim.x * blockDim.x)) {
    const int i = a.ix[lo + t], j = a.iy[lo + t], k = a.iz[lo + t];
    if (!inside(a.gb, i, j, k)) {
      atomicAdd(a.skipped, 1);
    } else if ((int)a.nb.at(i, j, k) <= a.mark_fb) {
      T r;
      double hh = -1.0e308;
      if (!inside(a.valid, i, j, k) || !local_curv_rhs<T>(a, i, j, k, r, hh)) {
        atomicAdd(a.skipped, 1);
      } else if (STAGE == 1) {
        a.rhs.at(i, j, k) = r;
        h = fmax(h, hh);
      } else {  // LSM3D_TVD_RK2_STAGE2_LOCAL
        a.out.at(i, j, k) = rk_combine<T>(RK2_STAGE2, a.pc.at(i, j, k), a.ps.at(i, j, k), r, *a.dt);
      }
    } else if (STAGE == 1) {
      a.rhs.at(i, j, k) = T(0);  // zeroed, and no term passes the mark
    }
  }
  if (STAGE == 1 && a.normal) {
    double c = 0.0;
    block_reduce2(RED_LIST_MAX, h, c);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = h;
  }
}

// LSM3D_TVD_RK2_STAGE1_LOCAL on level 0 with the device dt.
template <typename T>
__global__ void __launch_bounds__(256) k_local_curv_rk1(LocalCurvArgs<T> a) {
  const int lo = a.bs->n_lo[0], n = a.bs->n_hi[0] - lo + 1;
  for (int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); t < n; t += (int)(gridDim.x * blockDim.x)) {
    const int i = a.ix[lo + t], j = a.iy[lo + t], k = a.iz[lo + t];
    if (!inside(a.gb, i, j, k)) {
      atomicAdd(a.skipped, 1);
      continue;
    }
    if ((int)a.nb.at(i, j, k) > a.mark_fb) continue;
    const T u = a.ps.at(i, j, k);
    a.out.at(i, j, k) = rk_combine<T>(RK_EULER, u, u, a.rhs.at(i, j, k), *a.dt);
  }
}

// IMPOSE_MASK_LOCAL of the model (entries n_lo[0] .. n_hi[0] - 1,
// src/serial/lsm_macros.h:108-117): dest = max(mask, src).
template <typename T>
__global__ void __launch_bounds__(256) k_local_mask(const BandState* bs, const int* ix, const int* iy,
                                                    const int* iz, Field<T> dest, Field<const T> src,
                                                    Field<const T> mask, Box gb, int* skipped) {
  const int lo = bs->n_lo[0], n = bs->n_hi[0] - lo;
  for (int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); t < n; t += (int)(gridDim.x * blockDim.x)) {
    const int i = ix[lo + t], j = iy[lo + t], k = iz[lo + t];
    if (!inside(gb, i, j, k)) {
      atomicAdd(skipped, 1);
      continue;
    }
    const T m = mask.at(i, j, k), v = src.at(i, j, k);
    dest.at(i, j, k) = (m > v) ? m : v;
  }
}

// The step's dt: the CFL reduction's last step
// (LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL: cfl / (max H/dx + tiny) in
// the library precision), then the model's choice and clamps in double
// (curvature_model3d_local.c:331-349, 383-394).  dts[0] = dt, dtr[0] = (T) dt.
struct LocalDtArgs {
  const double* partials;
  int nblocks, normal, curv;
  double a, b, cfl, dx0, dx1, dx2, dt_sub, tplot, dt_min_to_correct, dt_min;
  BandState* bs;  // receives dt
};

template <typename T>
__global__ void __launch_bounds__(256) k_local_curv_dt(LocalDtArgs d, T cfl_r, T* dtr) {
#pragma clang fp contract(off)
  double v = -1.0e308, c = 0.0;
  for (int q = threadIdx.x; q < d.nblocks; q += blockDim.x) v = fmax(v, d.partials[q]);
  block_reduce2(RED_LIST_MAX, v, c);
  if (threadIdx.x != 0) return;
  double dt = d.tplot;
  if (d.normal) {
    T mx = T(-1.0);
    if (v > (double)mx) mx = T(v);
    dt = (double)(cfl_r / (mx + Tol<T>::tiny));
  }
  if (d.curv) {
    const double corr = (1.0 / (d.dx0 * d.dx0) + 1.0 / (d.dx1 * d.dx1) + 1.0 / (d.dx2 * d.dx2)) * (2.0 * d.b);
    const double max_h = d.normal ? d.cfl / dt : 0.0;
    dt = d.cfl / (max_h + corr);
  }
  if (dt < d.dt_min_to_correct) dt = d.dt_min;
  if (d.dt_sub + dt > d.tplot) dt = d.tplot - d.dt_sub;
  d.bs->dt = dt;
  dtr[0] = (T)dt;
}

}  // namespace lsm

// ============================================================================
// Host side
// ============================================================================
namespace {

using namespace lsm;
typedef real_t R;

#define BOXP(p)                                                                                \
  const int *ilo_##p, const int *ihi_##p, const int *jlo_##p, const int *jhi_##p, const int *klo_##p, \
      const int *khi_##p
#define BOX(p) lsm::box_of(ilo_##p, ihi_##p, jlo_##p, jhi_##p, klo_##p, khi_##p)

Box shrink(Box b, int hx, int hy, int hz) {
  return Box{b.ilo + hx, b.ihi - hx, b.jlo + hy, b.jhi - hy, b.klo + hz, b.khi - hz};
}
Box shrink_dir(Box b, int d, int lo, int hi) {
  if (d == 0) { b.ilo += lo; b.ihi -= hi; }
  else if (d == 1) { b.jlo += lo; b.jhi -= hi; }
  else { b.klo += lo; b.khi -= hi; }
  return b;
}
Box isect(Box a, const Box& b) {
  a.ilo = std::max(a.ilo, b.ilo); a.ihi = std::min(a.ihi, b.ihi);
  a.jlo = std::max(a.jlo, b.jlo); a.jhi = std::min(a.jhi, b.jhi);
  a.klo = std::max(a.klo, b.klo); a.khi = std::min(a.khi, b.khi);
  return a;
}

Field<const R> none() {
  Field<const R> f;
  f.p = nullptr;
  f.ilo = f.jlo = f.klo = 0;
  f.sy = f.sz = 0;
  return f;
}

int* skipped_counter() {
  bool fresh = false;
  int* p = reinterpret_cast<int*>(device_buffer(BUF_LOCAL_FLAGS, 64, &fresh));
  if (p && fresh) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, current_stream(), p, (int64_t)16, 0);
  }
  return p;
}

// The list arguments of one call, staged once with the largest extent used.
struct ListArgs {
  const int* ix = nullptr;
  const int* iy = nullptr;
  const int* iz = nullptr;
  int base = 0;  // the nlo the arrays are declared from
  bool ok = false;
};

ListArgs stage_lists(Call& c, const int* ix, const int* iy, const int* iz, int base, int last) {
  ListArgs L;
  L.base = base;
  if (last < base) return L;  // nothing is visited
  const int64_t n = (int64_t)last - base + 1;
  L.ix = c.in_buf(ix, n);
  L.iy = c.in_buf(iy, n);
  L.iz = c.in_buf(iz, n);
  L.ok = c.ok();
  return L;
}

// The mark test of a launch: narrow_band on its box (nb.p == nullptr: none)
// and the mark; `any` is what every entry touches besides narrow_band.
struct Marks {
  Field<const nb_t> nb;
  Box gnb;
  int mark;
};

Marks no_marks() {
  Marks m;
  m.nb.p = nullptr;
  m.nb.ilo = m.nb.jlo = m.nb.klo = 0;
  m.nb.sy = m.nb.sz = 0;
  m.gnb = Box{INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX};
  m.mark = 0;
  return m;
}

Entries make_entries(const ListArgs& L, int lo, int hi, const Marks& mk, Box valid, Box any) {
  Entries e;
  e.ix = L.ix + (lo - L.base);
  e.iy = L.iy + (lo - L.base);
  e.iz = L.iz + (lo - L.base);
  e.n = hi - lo + 1;
  e.nb = mk.nb;
  e.mark = mk.mark;
  e.any = isect(any, mk.gnb);
  e.valid = isect(valid, e.any);
  e.skipped = skipped_counter();
  return e;
}

Box everywhere() { return Box{INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX}; }

// Launches op over entries lo..hi of the staged lists.  Unmarked entries
// need only lie in `any` (and narrow_band's box); marked ones in `valid`.
template <typename Op>
void run_entries(Call& c, const char* kname, const ListArgs& L, int lo, int hi, const Marks& mk, Box valid, Op op,
                 Box any = everywhere()) {
  if (!c.ok() || !L.ok || hi < lo) return;
  if (mk.nb.p == nullptr) any = isect(any, valid);
  const int n = hi - lo + 1;
  Entries e = make_entries(L, lo, hi, mk, valid, any);
  if (!c.ok()) return;
  hipLaunchKernelGGL(k_entries<Op>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream(), e, op);
  c.launched(kname);
}

// Runs a list reduction; returns {value, count} (FP64).
template <typename Op>
void run_entries_reduce(Call& c, const char* kname, const ListArgs& L, int lo, int hi, const Marks& mk, Box valid,
                        Op op, int red, double out[2]) {
  out[0] = (red == RED_LIST_MAX) ? -1.0e308 : 0.0;
  out[1] = 0.0;
  if (!c.ok() || !L.ok || hi < lo) return;
  const int n = hi - lo + 1;
  const int nblocks = (n + 255) / 256;
  double* scratch = reinterpret_cast<double*>(device_scratch((size_t)(2 * nblocks + 2) * sizeof(double)));
  Entries e = make_entries(L, lo, hi, mk, valid, mk.nb.p ? everywhere() : valid);
  if (!c.ok()) return;
  hipLaunchKernelGGL(k_entries_reduce<Op>, dim3((unsigned)nblocks), dim3(256), 0, c.stream(), e, op, red,
                     scratch + 2);
  c.launched(kname);
  hipLaunchKernelGGL(k_entries_reduce_final, dim3(1), dim3(256), 0, c.stream(), scratch + 2, nblocks, red,
                     scratch);
  c.launched("k_entries_reduce_final");
  check_hip(hipMemcpyAsync(out, scratch, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream()),
            "hipMemcpyAsync(reduction result)");
  check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
}

// Completes a call; in synchronous mode (or when `force`) reports skipped
// list entries as an error.
void finish_local(Call& c, bool force = false) {
  int* sk = skipped_counter();
  if (!c.ok() || !sk || !(force || sync_on_return())) {
    c.finish();
    return;
  }
  int h = 0;
  check_hip(hipMemcpyAsync(&h, sk, sizeof(int), hipMemcpyDeviceToHost, c.stream()), "hipMemcpyAsync(flags)");
  c.finish();
  check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  if (h != 0) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c.stream(), sk, (int64_t)1, 0);
    (void)hipStreamSynchronize(c.stream());
    raise_error(std::string(c.name()) + ": " + std::to_string(h) +
                " list entries lie outside the arrays they index (skipped)");
  }
}

Marks marks(Call& c, const unsigned char* nb, const Box& b, const unsigned char* mark) {
  Marks m;
  m.nb = c.in(reinterpret_cast<const nb_t*>(nb), b);
  m.gnb = b;
  m.mark = (int)(nb_t)*mark;
  return m;
}

int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Band builders
// ---------------------------------------------------------------------------
bool host_array(const char* fn, const void* p, const char* what) {
  if (p == nullptr || is_device_pointer(p)) {
    raise_error(std::string(fn) + ": " + what + " must be host memory");
    return false;
  }
  return true;
}

// lsm3dMarkNarrowBandNeighbors (lsm_localization3d.f:48-168) on the staged
// device lists; `count` is the next free entry, updated.
void mark_neighbours(Call& c, const char* fn, Field<nb_t> nb, const Box& bn, int* dix, int* diy, int* diz,
                     int nlo, int nhi, int* n_lo, int* n_hi, int level) {
  const int64_t ncell = bn.cells();
  bool fresh = false;
  size_t got = 0;
  int* key = reinterpret_cast<int*>(device_buffer(BUF_BAND_KEY, (size_t)ncell * sizeof(int), &fresh, &got));
  if (!c.ok()) return;
  if (fresh) {  // the whole buffer holds INT_MAX between calls
    const int64_t nk = (int64_t)(got / sizeof(int));
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks_for(nk, 256)), dim3(256), 0, c.stream(), key, nk, INT_MAX);
  }
  CellMap cm{bn, bn.nx(), bn.nx() * bn.ny(), FastDiv((uint32_t)bn.nx()), FastDiv((uint32_t)(bn.nx() * bn.ny()))};
  Field<const nb_t> nbc = make_field<const nb_t>(nb.p, bn);
  int count = n_hi[0] + 1;
  for (int l = 1; l <= level && c.ok(); l++) {
    const int plo = n_lo[l - 1], phi_ = n_hi[l - 1];
    int added = 0;
    if (plo >= nlo && phi_ < count && phi_ >= plo) {
      const int64_t npairs64 = 6 * ((int64_t)phi_ - plo + 1);
      if (npairs64 > INT_MAX) {
        raise_error(std::string(fn) + ": narrow-band level too large");
        return;
      }
      const int npairs = (int)npairs64;
      const int nb_blocks = blocks_for(npairs, kPairChunk);
      unsigned char* work = reinterpret_cast<unsigned char*>(
          device_buffer(BUF_BAND_WORK, (size_t)npairs + (size_t)(3 * nb_blocks + 4) * sizeof(int) + 256));
      if (!c.ok()) return;
      int* counts = reinterpret_cast<int*>(work);
      int* totals = counts + 3 * nb_blocks;
      unsigned char* win = reinterpret_cast<unsigned char*>(totals + 4);
      Parents P{dix + (plo - nlo), diy + (plo - nlo), diz + (plo - nlo), npairs};
      hipLaunchKernelGGL(k_nbr_bid, dim3((unsigned)blocks_for(npairs, 256)), dim3(256), 0, c.stream(), P, nbc, cm,
                         key);
      c.launched("k_nbr_bid");
      hipLaunchKernelGGL(k_nbr_count, dim3((unsigned)nb_blocks), dim3(256), 0, c.stream(), P, nbc, cm, key, win,
                         counts);
      c.launched("k_nbr_count");
      hipLaunchKernelGGL(k_scan3, dim3(1), dim3(1024), 0, c.stream(), counts, nb_blocks, totals);
      c.launched("k_scan3");
      int tot[3] = {0, 0, 0};
      check_hip(hipMemcpyAsync(tot, totals, 3 * sizeof(int), hipMemcpyDeviceToHost, c.stream()),
                "hipMemcpyAsync(level size)");
      check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
      if (!c.ok()) return;
      const bool fits = (int64_t)count + tot[0] - 1 <= (int64_t)nhi;
      hipLaunchKernelGGL(k_nbr_write, dim3((unsigned)nb_blocks), dim3(256), 0, c.stream(), P, nb, cm, key, win,
                         counts, dix + (count - nlo), diy + (count - nlo), diz + (count - nlo), nb_t(l + 1),
                         fits ? 1 : 0);
      c.launched("k_nbr_write");
      if (!fits) {
        raise_error(std::string(fn) + ": the index arrays are too short for the narrow band");
        return;
      }
      added = tot[0];
    }
    count += added;
    if (count > n_hi[l - 1] + 1) {
      n_lo[l] = n_hi[l - 1] + 1;
      n_hi[l] = count - 1;
    } else {
      n_lo[l] = -1;
      n_hi[l] = -1;
    }
  }
}

void determine_band(const char* fn, int mode, const R* phi, Box g, unsigned char* nb_arr, Box bn, int* ix,
                    int* iy, int* iz, int nlo, int nhi, int* n_lo, int* n_hi, int* outer, int nlo_o, int nhi_o,
                    int* nlo_plus, int* nhi_plus, int* nlo_minus, int* nhi_minus, R width, R width_inner,
                    int level) {
  Call c(fn);
  if (!host_array(fn, n_lo, "n_lo") || !host_array(fn, n_hi, "n_hi")) return;
  if (level < 0) {
    raise_error(std::string(fn) + ": level must be >= 0");
    return;
  }
  require_inside(fn, "narrow_band", bn, g, 0, 0, 0);
  if (g.cells() <= 0 || g.cells() > INT_MAX || bn.cells() > INT_MAX) {
    raise_error(std::string(fn) + ": box sizes must be in 1 .. 2^31-1 cells");
    return;
  }
  const int64_t cap = (int64_t)nhi - nlo + 1;
  auto fphi = c.in(phi, g);
  Field<nb_t> fnb = c.out(reinterpret_cast<nb_t*>(nb_arr), bn);
  int* dix = c.out_buf(ix, cap > 0 ? cap : 1);
  int* diy = c.out_buf(iy, cap > 0 ? cap : 1);
  int* diz = c.out_buf(iz, cap > 0 ? cap : 1);
  const int64_t ocap = (int64_t)nhi_o - nlo_o + 1;
  int* dout = nullptr;
  if (mode == BAND_PHI) dout = c.out_buf(outer, ocap > 0 ? ocap : 1);
  if (!c.ok()) return;

  const int ncell = (int)g.cells();
  const int nblk = blocks_for(ncell, kChunk);
  int* work = reinterpret_cast<int*>(device_buffer(BUF_BAND_WORK, (size_t)(3 * nblk + 4) * sizeof(int) + 256));
  if (!c.ok()) return;
  int* totals = work + 3 * nblk;
  CellMap cm{g, g.nx(), g.nx() * g.ny(), FastDiv((uint32_t)g.nx()), FastDiv((uint32_t)(g.nx() * g.ny()))};
  const int same_box = (bn.ilo == g.ilo && bn.ihi == g.ihi && bn.jlo == g.jlo && bn.jhi == g.jhi &&
                        bn.klo == g.klo && bn.khi == g.khi) ? 1 : 0;
  // narrow_band on phi's box: the count pass stores each cell's flags there and
  // the write pass reads them back instead of phi
  nb_t* codes = same_box ? fnb.p : nullptr;
  hipLaunchKernelGGL(k_band_count<R>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), fphi, ncell, mode, width,
                     width_inner, work, codes);
  c.launched("k_band_count");
  hipLaunchKernelGGL(k_scan3, dim3(1), dim3(1024), 0, c.stream(), work, nblk, totals);
  c.launched("k_scan3");
  int tot[3] = {0, 0, 0};
  check_hip(hipMemcpyAsync(tot, totals, 3 * sizeof(int), hipMemcpyDeviceToHost, c.stream()),
            "hipMemcpyAsync(band size)");
  check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  if (!c.ok()) return;
  if (tot[0] > cap) {
    raise_error(std::string(fn) + ": the index arrays are too short for the narrow band");
    return;
  }
  if (mode == BAND_PHI && (int64_t)tot[1] + tot[2] > ocap) {
    raise_error(std::string(fn) + ": index_outer is too short for the outer layer");
    return;
  }
  if (codes) {
    if (((uintptr_t)codes & 15) == 0)
      hipLaunchKernelGGL(k_band_write_codes<true>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), codes, cm, ncell,
                         work, dix, diy, diz, nlo, dout, (int)(ocap - 1), (const int*)nullptr);
    else
      hipLaunchKernelGGL(k_band_write_codes<false>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), codes, cm,
                         ncell, work, dix, diy, diz, nlo, dout, (int)(ocap - 1), (const int*)nullptr);
    c.launched("k_band_write_codes");
  } else {
    hipLaunchKernelGGL(k_band_write<R>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), fphi, cm, ncell, mode,
                       width, width_inner, work, fnb, same_box, dix, diy, diz, nlo, dout, (int)(ocap - 1));
    c.launched("k_band_write");
  }
  if (!c.ok()) return;

  const int count = nlo + tot[0];
  n_lo[0] = nlo;
  if (mode == BAND_PHI) {
    *nlo_minus = nlo_o;
    *nhi_plus = nhi_o;
    if (count > nlo) {
      n_hi[0] = count - 1;
      *nhi_minus = nlo_o + tot[1] - 1;
      *nlo_plus = nhi_o - tot[2] + 1;
      mark_neighbours(c, fn, fnb, bn, dix, diy, diz, nlo, nhi, n_lo, n_hi, level);
    } else {
      n_hi[0] = count;
      *nhi_minus = nlo_o;
      *nlo_plus = nhi_o;
    }
  } else {
    n_hi[0] = count - 1;
    mark_neighbours(c, fn, fnb, bn, dix, diy, diz, nlo, nhi, n_lo, n_hi, level);
  }
  c.finish();
}

// Shared entry for the TVD Runge-Kutta stages on the band.
void rk_local(const char* fn, int form, R* out, Box go, const R* us, Box gs, const R* uc, Box gc, const R* rhs,
              Box gr, R dt, const int* ix, const int* iy, const int* iz, int nlo, int nhi,
              const unsigned char* nb, Box gnb, const unsigned char* mark) {
  Call c(fn);
  ListArgs L = stage_lists(c, ix, iy, iz, nlo, nhi);
  if (L.ok) {
    Box valid = isect(isect(isect(go, gc), gr), gnb);
    if (us) valid = isect(valid, gs);
    OpRk<R> op{form, c.out(out, go), us ? c.in(us, gs) : none(), c.in(uc, gc), c.in(rhs, gr), dt};
    run_entries(c, "k_entries<rk>", L, nlo, nhi, marks(c, nb, gnb, mark), valid, op);
  }
  finish_local(c);
}

R max3(R a, R b, R c) {
  R m = a;
  if (b > m) m = b;
  if (c > m) m = c;
  return m;
}

// lsm3dHJENO1LOCAL / lsm3dHJENO2LOCAL: the D arrays are written as the
// reference writes them (the last direction's differences remain).
void hj_eno_local(const char* fn, int order, R* xp, R* yp, R* zp, Box gp, R* xm, R* ym, R* zm, Box gm,
                  const R* phi, Box gphi, R* D1, Box g1, R* D2, Box g2, R dx, R dy, R dz, const int* ix,
                  const int* iy, const int* iz, int nlo0, int nhi0, int nhi1, int nhi2, const unsigned char* nb,
                  Box gnb, const unsigned char* mark_fb, const unsigned char* mark_d1,
                  const unsigned char* mark_d2) {
  Call c(fn);
  const int d1_hi = (order == 1) ? nhi1 : nhi2;  // D1 over levels 0..1 (ENO1) or 0..2 (ENO2)
  const int last = std::max(nhi0, std::max(d1_hi, order == 2 ? nhi1 : nhi0));
  ListArgs L = stage_lists(c, ix, iy, iz, nlo0, last);
  if (L.ok) {
    Field<R> fxp = c.out(xp, gp), fyp = c.out(yp, gp), fzp = c.out(zp, gp);
    Field<R> fxm = c.out(xm, gm), fym = c.out(ym, gm), fzm = c.out(zm, gm);
    Field<const R> fphi = c.in(phi, gphi);
    Field<R> f1 = c.out(D1, g1);
    Field<R> f2 = (order == 2) ? c.out(D2, g2) : Field<R>{};
    const Marks mfb = marks(c, nb, gnb, mark_fb), m1 = marks(c, nb, gnb, mark_d1);
    const Marks m2 = (order == 2) ? marks(c, nb, gnb, mark_d2) : no_marks();
    Field<const R> c1 = make_field<const R>(f1.p, g1), c2 = make_field<const R>(f2.p, g2);
    Field<R> outp[3] = {fxp, fyp, fzp}, outm[3] = {fxm, fym, fzm};
    const R inv[3] = {R(1.0 / (double)dx), R(1.0 / (double)dy), R(1.0 / (double)dz)};
    for (int d = 0; d < 3 && c.ok(); d++) {
      const int e[3] = {d == 0, d == 1, d == 2};
      // D1(c) = phi(c) - phi(c - e_d) over levels 0..d1_hi (1e10 where unmarked)
      run_entries(c, "k_entries<dn1>", L, nlo0, d1_hi, m1, isect(g1, shrink_dir(gphi, d, 1, 0)),
                  OpDn<R>{f1, fphi, e[0], e[1], e[2], R(1)}, g1);
      if (order == 2) {
        // D2(c) = -(D1(c) - D1(c + e_d)) over levels 0..1
        run_entries(c, "k_entries<dn2>", L, nlo0, nhi1, m2, isect(g2, shrink_dir(g1, d, 0, 1)),
                    OpDn<R>{f2, c1, -e[0], -e[1], -e[2], R(-1)}, g2);
        Box v = isect(isect(isect(gp, gm), shrink_dir(g2, d, 1, 1)), shrink_dir(g1, d, 0, 1));
        run_entries(c, "k_entries<eno2>", L, nlo0, nhi0, mfb, v,
                    OpEno2<R>{outp[d], outm[d], c1, c2, e[0], e[1], e[2], inv[d]});
      } else {
        Box v = isect(isect(gp, gm), shrink_dir(g1, d, 0, 1));
        run_entries(c, "k_entries<eno1>", L, nlo0, nhi0, mfb, v,
                    OpEno1<R>{outp[d], outm[d], c1, e[0], e[1], e[2], inv[d]});
      }
    }
  }
  finish_local(c);
}

}  // namespace

// ============================================================================
// The narrow-band curvature model's fused stages (lsm_local_step.h)
// ============================================================================
namespace lsm {
namespace {

// Workgroups of the launches sized by the band state (they stride over
// their entries, pairs or chunks).
constexpr int kLocalGrid = 2048;
constexpr int kStateBytes = 512;  // BandState, then the library-precision dt at byte 256

// The step's device memory: the band state, the library-precision dt, and
// one CFL maximum per workgroup of the stage-1 launch.  Every call of a
// step asks for the same size, so the buffer does not move between them.
char* step_buffer() {
  return reinterpret_cast<char*>(
      device_buffer(BUF_LOCAL_STEP, (size_t)kStateBytes + (size_t)kLocalGrid * sizeof(double)));
}
BandState* band_state(char* buf) { return reinterpret_cast<BandState*>(buf); }
R* step_dtr(char* buf) { return reinterpret_cast<R*>(buf + 256); }
double* step_partials(char* buf) { return reinterpret_cast<double*>(buf + kStateBytes); }

LocalCurvArgs<R> local_args(Call& c, const LocalCurvStep& s, const R* ps, const R* pc, R* rhs, R* out,
                            char* buf) {
  LocalCurvArgs<R> a;
  a.ix = s.index_x;
  a.iy = s.index_y;
  a.iz = s.index_z;
  a.bs = band_state(buf);
  a.nb = c.in(reinterpret_cast<const nb_t*>(s.narrow_band), s.gb);
  a.ps = c.in(ps, s.gb);
  a.pc = pc ? c.in(pc, s.gb) : none();
  a.rhs = make_field<R>(rhs, s.gb);
  a.out = make_field<R>(out, s.gb);
  a.gb = s.gb;
  a.valid = shrink(s.gb, 2, 2, 2);
  a.mark_fb = (int)(nb_t)s.mark_fb;
  a.mark_d1 = (int)(nb_t)s.mark_d1;
  a.mark_d2 = (int)(nb_t)s.mark_d2;
  a.normal = s.a > 0;
  a.curv = s.b > 0;
  a.vn = (R)s.a;
  a.avn = R(fabs(a.vn));
  a.b = (R)s.b;
  const R dx = (R)s.dx[0], dy = (R)s.dx[1], dz = (R)s.dx[2];
  a.inv[0] = R(1.0 / (double)dx);
  a.inv[1] = R(1.0 / (double)dy);
  a.inv[2] = R(1.0 / (double)dz);
  a.half[0] = R(0.5 / (double)dx);
  a.half[1] = R(0.5 / (double)dy);
  a.half[2] = R(0.5 / (double)dz);
  const R mdx = max3(dx, dy, dz);
  a.mdx2 = mdx * mdx;
  const R beta = (R)(2.0 * s.dx[0]), gamma = (R)(4.0 * s.dx[0]);  // the model's band widths
  a.beta = beta;
  a.gamma = gamma;
  a.c1 = gamma - R(3) * beta;
  R c2 = gamma - beta;
  a.c2 = c2 * c2 * c2;
  a.dt = step_dtr(buf);
  a.partials = step_partials(buf);
  a.skipped = skipped_counter();
  return a;
}

bool step_ok(Call& c, char* buf, const char* what) {
  if (!buf) raise_error(std::string("lsm3d_amd_local_curvature_step: ") + what + ": device buffer allocation failed");
  return c.ok();
}

}  // namespace

void local_band(const LocalCurvStep& s) {
  Call c("lsm3d_amd_local_curvature_step (band)");
  char* buf = step_buffer();
  if (!step_ok(c, buf, "band")) return;
  const Box g = s.gb;
  const int64_t ncell64 = g.cells();
  if (ncell64 <= 0 || ncell64 > INT_MAX || s.num_points < ncell64 || s.levels < 0 || s.levels > 7) {
    raise_error("lsm3d_amd_local_curvature_step: the box, the list capacity or the level count is out of range");
    return;
  }
  const int ncell = (int)ncell64;
  const R width = (R)(4.0 * s.dx[0]), width_inner = (R)(2.0 * s.dx[0]);  // gamma, beta of the model
  Field<const R> fphi = c.in(s.phi, g);
  Field<nb_t> fnb = c.out(reinterpret_cast<nb_t*>(s.narrow_band), g);
  const int nlo = 0, nhi = s.num_points - 1;
  const int nblk = blocks_for(ncell, kChunk);
  // level 0's block counts, then the largest neighbour level's chunk counts
  const int64_t max_chunks = (6 * (int64_t)s.num_points + kPairChunk - 1) / kPairChunk;
  const int64_t rows = std::max<int64_t>(nblk, max_chunks);
  int* work = reinterpret_cast<int*>(device_buffer(BUF_BAND_WORK, (size_t)(3 * rows + 8) * sizeof(int) + 256));
  bool fresh = false;
  size_t got = 0;
  int* key = reinterpret_cast<int*>(device_buffer(BUF_BAND_KEY, (size_t)ncell * sizeof(int), &fresh, &got));
  if (!work || !key) raise_error("lsm3d_amd_local_curvature_step: band buffers: device allocation failed");
  if (!c.ok()) return;
  if (fresh) {  // the whole key buffer holds INT_MAX between calls
    const int64_t nk = (int64_t)(got / sizeof(int));
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks_for(nk, 256)), dim3(256), 0, c.stream(), key, nk, INT_MAX);
  }
  int* totals = work + 3 * rows;
  BandState* bs = band_state(buf);
  CellMap cm{g, g.nx(), g.nx() * g.ny(), FastDiv((uint32_t)g.nx()), FastDiv((uint32_t)(g.nx() * g.ny()))};
  // level 0: flags in narrow_band, scan, lists (LSM3D_DETERMINE_NARROW_BAND's scan, lsm_localization3d.f:232-255)
  hipLaunchKernelGGL(k_band_count<R>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), fphi, ncell, (int)BAND_PHI,
                     width, width_inner, work, fnb.p);
  c.launched("k_band_count");
  hipLaunchKernelGGL(k_scan3, dim3(1), dim3(1024), 0, c.stream(), work, nblk, totals);
  c.launched("k_scan3");
  BandPrev prev;
  for (int l = 0; l < 8; l++) {
    prev.lo[l] = s.prev_lo[l];
    prev.hi[l] = s.prev_hi[l];
  }
  hipLaunchKernelGGL(k_band_level0, dim3(1), dim3(1), 0, c.stream(), bs, (const int*)totals, nlo, s.num_points,
                     s.num_points, s.levels, prev);
  c.launched("k_band_level0");
  if (((uintptr_t)fnb.p & 15) == 0)
    hipLaunchKernelGGL(k_band_write_codes<true>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), fnb.p, cm, ncell,
                       (const int*)work, s.index_x, s.index_y, s.index_z, nlo, s.index_outer, s.num_points - 1,
                       (const int*)&bs->err);
  else
    hipLaunchKernelGGL(k_band_write_codes<false>, dim3((unsigned)nblk), dim3(256), 0, c.stream(), fnb.p, cm, ncell,
                       (const int*)work, s.index_x, s.index_y, s.index_z, nlo, s.index_outer, s.num_points - 1,
                       (const int*)&bs->err);
  c.launched("k_band_write_codes");
  // levels 1 .. levels: first-discovery bids, winners compacted in pair order
  // (lsm3dMarkNarrowBandNeighbors, lsm_localization3d.f:48-168)
  const Field<const nb_t> nbc = make_field<const nb_t>(fnb.p, g);
  for (int l = 1; l <= s.levels && c.ok(); l++) {
    hipLaunchKernelGGL(k_band_level_setup, dim3(1), dim3(1), 0, c.stream(), bs, l, nlo);
    hipLaunchKernelGGL(k_nbr_bid_dev, dim3(4 * kLocalGrid), dim3(256), 0, c.stream(), (const BandState*)bs,
                       (const int*)s.index_x, (const int*)s.index_y, (const int*)s.index_z, nlo, nbc, cm, key);
    hipLaunchKernelGGL(k_nbr_count_dev, dim3(kLocalGrid), dim3(256), 0, c.stream(), (const BandState*)bs,
                       (const int*)s.index_x, (const int*)s.index_y, (const int*)s.index_z, nlo, nbc, cm,
                       (const int*)key, work);
    hipLaunchKernelGGL(k_scan3_dev, dim3(1), dim3(1024), 0, c.stream(), work, (const int*)&bs->nchunks, totals);
    hipLaunchKernelGGL(k_nbr_write_dev, dim3(kLocalGrid), dim3(256), 0, c.stream(), bs, (const int*)totals,
                       (const int*)work, s.index_x, s.index_y, s.index_z, nlo, nhi, fnb, cm, key, nb_t(l + 1));
    hipLaunchKernelGGL(k_band_level_finish, dim3(1), dim3(1), 0, c.stream(), bs, (const int*)totals, l);
    c.launched("neighbour level");
  }
  c.finish();
}

void local_curv_stage1(const LocalCurvStep& s, double dt_sub, double tplot) {
  Call c("lsm3d_amd_local_curvature_step (stage 1)");
  char* buf = step_buffer();
  if (!step_ok(c, buf, "stage 1")) return;
  LocalCurvArgs<R> a = local_args(c, s, s.phi, nullptr, s.lse_rhs, s.phi_stage1, buf);
  if (!c.ok()) return;
  hipLaunchKernelGGL((k_local_curv_stage<R, 1>), dim3(kLocalGrid), dim3(256), 0, c.stream(), a);
  c.launched("k_local_curv_stage<1>");
  LocalDtArgs d{step_partials(buf), a.normal ? kLocalGrid : 0, a.normal, a.curv, s.a, s.b, s.cfl, s.dx[0],
                s.dx[1], s.dx[2], dt_sub, tplot, s.dt_min_to_correct, s.dt_min, band_state(buf)};
  hipLaunchKernelGGL(k_local_curv_dt<R>, dim3(1), dim3(256), 0, c.stream(), d, (R)s.cfl, step_dtr(buf));
  c.launched("k_local_curv_dt");
  finish_local(c);
}

void local_curv_rk1(const LocalCurvStep& s) {
  Call c("lsm3d_amd_local_curvature_step (RK2 stage 1)");
  char* buf = step_buffer();
  if (!step_ok(c, buf, "RK2 stage 1")) return;
  LocalCurvArgs<R> a = local_args(c, s, s.phi, nullptr, s.lse_rhs, s.phi_stage1, buf);
  if (!c.ok()) return;
  hipLaunchKernelGGL(k_local_curv_rk1<R>, dim3(kLocalGrid), dim3(256), 0, c.stream(), a);
  c.launched("k_local_curv_rk1");
  finish_local(c);
}

void local_curv_stage2(const LocalCurvStep& s) {
  Call c("lsm3d_amd_local_curvature_step (stage 2)");
  char* buf = step_buffer();
  if (!step_ok(c, buf, "stage 2")) return;
  LocalCurvArgs<R> a = local_args(c, s, s.phi_stage1, s.phi, s.lse_rhs, s.phi_next, buf);
  if (!c.ok()) return;
  hipLaunchKernelGGL((k_local_curv_stage<R, 2>), dim3(kLocalGrid), dim3(256), 0, c.stream(), a);
  c.launched("k_local_curv_stage<2>");
  finish_local(c);
}

void local_mask(const LocalCurvStep& s, R* dest, const R* src) {
  Call c("lsm3d_amd_local_curvature_step (mask)");
  char* buf = step_buffer();
  if (!step_ok(c, buf, "mask")) return;
  Field<R> fd = c.out(dest, s.gb);
  Field<const R> fs = c.in(src, s.gb), fm = c.in(s.mask, s.gb);
  if (!c.ok()) return;
  hipLaunchKernelGGL(k_local_mask<R>, dim3(kLocalGrid), dim3(256), 0, c.stream(), (const BandState*)band_state(buf),
                     (const int*)s.index_x, (const int*)s.index_y, (const int*)s.index_z, fd, fs, fm, s.gb,
                     skipped_counter());
  c.launched("k_local_mask");
  finish_local(c);
}

bool local_trigger_values(const LocalCurvStep& s, int nlo0, int nhi0, const int outer_limits[4], int* change,
                          R* ave) {
  Call c("lsm3d_amd_local_curvature_reinit_trigger");
  // LSM3D_CHECK_OUTER_NARROW_BAND_LAYER over the plus and minus outer lists
  // (entries resolved through level 0), then LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL
  // over level 0 with the fill-box mark: the routines' kernels and launch
  // shapes, read back together
  const int lo[2] = {outer_limits[0], outer_limits[2]}, hi[2] = {outer_limits[1], outer_limits[3]};
  const int last = s.num_points - 1;
  for (int q = 0; q < 2; q++)
    if (hi[q] >= lo[q] && (lo[q] < 0 || hi[q] > last)) {
      raise_error("lsm3d_amd_local_curvature_reinit_trigger: outer-layer range outside index_outer");
      return false;
    }
  ListArgs L = stage_lists(c, s.index_x, s.index_y, s.index_z, nlo0, nhi0);
  Field<const R> fphi = c.in(s.phi, s.gb);
  int* bits = skipped_counter();
  if (!c.ok() || !bits) return false;
  int* dbits = bits + 4;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c.stream(), dbits, (int64_t)1, 0);
  for (int q = 0; q < 2 && L.ok; q++) {
    if (hi[q] < lo[q]) continue;
    const int n = hi[q] - lo[q] + 1;
    hipLaunchKernelGGL(k_check_outer<R>, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, c.stream(),
                       (const int*)s.index_outer + lo[q], n, L.ix, L.iy, L.iz, nlo0, nhi0, fphi, s.gb, 2 * q, dbits,
                       bits);
    c.launched("k_check_outer");
  }
  double out[2] = {0.0, 0.0};
  double* scratch = nullptr;
  if (L.ok && nhi0 >= nlo0) {
    const int n = nhi0 - nlo0 + 1, nblocks = (n + 255) / 256;
    scratch = reinterpret_cast<double*>(device_scratch((size_t)(2 * nblocks + 2) * sizeof(double)));
    if (!scratch) raise_error("lsm3d_amd_local_curvature_reinit_trigger: device scratch allocation failed");
    if (!c.ok()) return false;
    const R dx = (R)s.dx[0], dy = (R)s.dx[1], dz = (R)s.dx[2];
    RedAveGrad<R> op{fphi, R(0.5 / (double)dx), R(0.5 / (double)dy), R(0.5 / (double)dz)};
    const unsigned char mfb = (unsigned char)s.mark_fb;
    const Marks mk = marks(c, s.narrow_band, s.gb, &mfb);
    Entries e = make_entries(L, nlo0, nhi0, mk, isect(shrink(s.gb, 1, 1, 1), s.gb), everywhere());
    hipLaunchKernelGGL(k_entries_reduce<RedAveGrad<R>>, dim3((unsigned)nblocks), dim3(256), 0, c.stream(), e, op,
                       (int)RED_LIST_SUM, scratch + 2);
    c.launched("k_entries_reduce<avegrad>");
    hipLaunchKernelGGL(k_entries_reduce_final, dim3(1), dim3(256), 0, c.stream(), (const double*)(scratch + 2),
                       nblocks, (int)RED_LIST_SUM, scratch);
    c.launched("k_entries_reduce_final");
  }
  int h = 0;
  if (c.ok()) {
    check_hip(hipMemcpyAsync(&h, dbits, sizeof(int), hipMemcpyDeviceToHost, c.stream()), "hipMemcpyAsync(bits)");
    if (scratch)
      check_hip(hipMemcpyAsync(out, scratch, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream()),
                "hipMemcpyAsync(average gradient)");
    check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  }
  finish_local(c, true);
  if (error_pending()) return false;
  *change = ((h & 3) == 3 || ((h >> 2) & 3) == 3) ? 1 : 0;
  *ave = (out[1] > 0.0) ? R(out[0] / out[1]) : R(out[0]);
  return true;
}

bool local_step_result(const LocalCurvStep& s, LocalStepResult* r) {
  (void)s;
  Call c("lsm3d_amd_local_curvature_step (result)");
  char* buf = step_buffer();
  if (!buf) raise_error("lsm3d_amd_local_curvature_step: result: device buffer allocation failed");
  BandState h;
  if (buf && c.ok()) {
    check_hip(hipMemcpyAsync(&h, buf, sizeof(BandState), hipMemcpyDeviceToHost, c.stream()),
              "hipMemcpyAsync(band state)");
    check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  }
  finish_local(c);
  if (error_pending()) return false;
  if (h.err) {
    raise_error("lsm3d_amd_local_curvature_step: the index arrays are too short for the narrow band");
    return false;
  }
  for (int l = 0; l < 8; l++) {
    r->n_lo[l] = h.n_lo[l];
    r->n_hi[l] = h.n_hi[l];
  }
  for (int q = 0; q < 3; q++) r->tot[q] = h.tot[q];
  r->dt = h.dt;
  return true;
}

}  // namespace lsm

// ============================================================================
// Entry points (extern "C", reference prototypes; see include/lsm3d_amd.h)
// ============================================================================
extern "C" {

// ---- narrow band: lsm_localization3d.h --------------------------------------
void lsm3ddeterminenarrowband_(const R* phi, BOXP(g), unsigned char* nb, BOXP(nb), int* ix, int* iy, int* iz,
                               const int* nlo_index, const int* nhi_index, int* n_lo, int* n_hi, int* index_outer,
                               const int* nlo_index_outer, const int* nhi_index_outer, int* nlo_outer_plus,
                               int* nhi_outer_plus, int* nlo_outer_minus, int* nhi_outer_minus, const R* width,
                               const R* width_inner, const int* level) {
  determine_band("LSM3D_DETERMINE_NARROW_BAND", BAND_PHI, phi, BOX(g), nb, BOX(nb), ix, iy, iz, *nlo_index,
                 *nhi_index, n_lo, n_hi, index_outer, *nlo_index_outer, *nhi_index_outer, nlo_outer_plus,
                 nhi_outer_plus, nlo_outer_minus, nhi_outer_minus, *width, *width_inner, *level);
}

void lsm3ddeterminenarrowbandfrommask_(const R* mask, BOXP(m), unsigned char* nb, BOXP(nb), int* ix, int* iy,
                                       int* iz, int* nlo_index, int* nhi_index, int* n_lo, int* n_hi,
                                       const int* level, const int* use_mask_sign) {
  determine_band("LSM3D_DETERMINE_NARROW_BAND_FROM_MASK",
                 *use_mask_sign > 0 ? BAND_MASK_NONNEG : BAND_MASK_NONPOS, mask, BOX(m), nb, BOX(nb), ix, iy, iz,
                 *nlo_index, *nhi_index, n_lo, n_hi, nullptr, 0, -1, nullptr, nullptr, nullptr, nullptr, R(0),
                 R(0), *level);
}

void lsm3dmarknarrowbandboundarylayer_(unsigned char* nb, BOXP(gb), BOXP(fb), const unsigned char* mark) {
  const char* fn = "LSM3D_MARK_NARROW_BAND_BOUNDARY_LAYER";
  Call c(fn);
  const Box gb = BOX(gb), fb = BOX(fb);
  Field<nb_t> f = c.out(reinterpret_cast<nb_t*>(nb), gb);
  if (c.ok()) {
    const int n = std::max(gb.nx(), std::max(gb.ny(), gb.nz()));
    hipLaunchKernelGGL(k_mark_layer, dim3((unsigned)((n + 63) / 64), (unsigned)((n + 3) / 4), 6), dim3(256), 0,
                       c.stream(), f, gb, fb, (nb_t)*mark);
    c.launched("k_mark_layer");
  }
  c.finish();
}

void lsm3dmultiplycutofflserhslocal_(const R* phi, R* rhs, BOXP(g), const int* ix, const int* iy, const int* iz,
                                     const int* nlo, const int* nhi, const unsigned char* nb, BOXP(nb),
                                     const unsigned char* mark_fb, const R* beta, const R* gamma) {
  Call c("LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box g = BOX(g), gnb = BOX(nb);
    const R b = *beta, gm = *gamma;
    R c2 = gm - b;
    c2 = c2 * c2 * c2;
    OpCutoff<R> op{c.out(rhs, g), c.in(phi, g), b, gm, gm - R(3) * b, c2};
    run_entries(c, "k_entries<cutoff>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb), isect(g, gnb), op);
  }
  finish_local(c);
}

void lsm3dcheckouternarrowbandlayer_(int* change_sign, const R* phi, BOXP(g), int* ix, int* iy,
                                     int* iz, int* nlo, int* nhi, const int* index_outer,
                                     const int* nlo_outer, const int* nhi_outer, const int* nlo_plus,
                                     const int* nhi_plus, const int* nlo_minus, const int* nhi_minus) {
  const char* fn = "LSM3D_CHECK_OUTER_NARROW_BAND_LAYER";
  Call c(fn);
  if (!host_array(fn, change_sign, "change_sign")) return;
  const Box g = BOX(g);
  const int lo[2] = {*nlo_plus, *nlo_minus}, hi[2] = {*nhi_plus, *nhi_minus};
  for (int s = 0; s < 2; s++)
    if (hi[s] >= lo[s] && (lo[s] < *nlo_outer || hi[s] > *nhi_outer)) {
      raise_error(std::string(fn) + ": outer-layer range outside index_outer");
      return;
    }
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  const int64_t ocap = (int64_t)*nhi_outer - *nlo_outer + 1;
  const int* dout = (ocap > 0 && (hi[0] >= lo[0] || hi[1] >= lo[1])) ? c.in_buf(index_outer, ocap) : nullptr;
  Field<const R> fphi = c.in(phi, g);
  int* bits = skipped_counter();
  if (!c.ok()) return;
  int* dbits = bits + 4;  // bits[4]: the sign bits of this call
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c.stream(), dbits, (int64_t)1, 0);
  for (int s = 0; s < 2 && dout && L.ok; s++) {
    if (hi[s] < lo[s]) continue;
    const int n = hi[s] - lo[s] + 1;
    hipLaunchKernelGGL(k_check_outer<R>, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, c.stream(),
                       dout + (lo[s] - *nlo_outer), n, L.ix, L.iy, L.iz, *nlo, *nhi, fphi, g, 2 * s, dbits, bits);
    c.launched("k_check_outer");
  }
  int h = 0;
  if (c.ok()) {
    check_hip(hipMemcpyAsync(&h, dbits, sizeof(int), hipMemcpyDeviceToHost, c.stream()), "hipMemcpyAsync(bits)");
    check_hip(hipStreamSynchronize(c.stream()), "hipStreamSynchronize");
  }
  finish_local(c, true);
  if (!lsm::error_pending()) *change_sign = ((h & 3) == 3 || ((h >> 2) & 3) == 3) ? 1 : 0;
}

void lsm3dimposemasklocal_(R* dest, BOXP(d), R* src, BOXP(s), R* mask, BOXP(m), const int* ix,
                           const int* iy, const int* iz, const int* nlo, const int* nhi) {
  Call c("LSM3D_IMPOSE_MASK_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gd = BOX(d), gs = BOX(s), gm = BOX(m);
    OpMaskMax<R> op{c.out(dest, gd), c.in(src, gs), c.in(mask, gm)};
    run_entries(c, "k_entries<mask>", L, *nlo, *nhi, no_marks(), isect(isect(gd, gs), gm), op);
  }
  finish_local(c);
}

void lsm3dcopydatalocal_(const R* dest_out, BOXP(d), const R* src, BOXP(s), const int* ix, const int* iy,
                         const int* iz, const int* nlo, const int* nhi) {
  R* dest = const_cast<R*>(dest_out);  // output, declared const by the reference
  Call c("LSM3D_COPY_DATA_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gd = BOX(d), gs = BOX(s);
    OpCopy<R> op{c.out(dest, gd), c.in(src, gs)};
    run_entries(c, "k_entries<copy>", L, *nlo, *nhi, no_marks(), isect(gd, gs), op);
  }
  finish_local(c);
}

// ---- spatial derivatives: lsm_spatial_derivatives3d_local.h -------------------
void lsm3dhjeno1local_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                       R* D1, BOXP(d1), const R* dx, const R* dy, const R* dz, const int* ix, const int* iy,
                       const int* iz, const int* nlo0, const int* nhi0, const int* nlo1, const int* nhi1,
                       const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb,
                       const unsigned char* mark_d1) {
  (void)nlo1;
  hj_eno_local("LSM3D_HJ_ENO1_LOCAL", 1, xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), D1, BOX(d1),
               nullptr, Box{}, *dx, *dy, *dz, ix, iy, iz, *nlo0, *nhi0, *nhi1, *nhi1, nb, BOX(nb), mark_fb,
               mark_d1, nullptr);
}

void lsm3dhjeno2local_(R* xp, R* yp, R* zp, BOXP(gp), R* xm, R* ym, R* zm, BOXP(gm), const R* phi, BOXP(phi),
                       R* D1, BOXP(d1), R* D2, BOXP(d2), const R* dx, const R* dy, const R* dz, const int* ix,
                       const int* iy, const int* iz, const int* nlo0, const int* nhi0, const int* nlo1,
                       const int* nhi1, const int* nlo2, const int* nhi2, const unsigned char* nb, BOXP(nb),
                       const unsigned char* mark_fb, const unsigned char* mark_d1, const unsigned char* mark_d2) {
  (void)nlo1;
  (void)nlo2;
  hj_eno_local("LSM3D_HJ_ENO2_LOCAL", 2, xp, yp, zp, BOX(gp), xm, ym, zm, BOX(gm), phi, BOX(phi), D1, BOX(d1), D2,
               BOX(d2), *dx, *dy, *dz, ix, iy, iz, *nlo0, *nhi0, *nhi1, *nhi2, nb, BOX(nb), mark_fb, mark_d1,
               mark_d2);
}

static void central_local(const char* fn, int order, R* gx, R* gy, R* gz, Box gg, const R* phi, Box gphi, R dx,
                          R dy, R dz, const int* ix, const int* iy, const int* iz, int nlo, int nhi,
                          const unsigned char* nb, Box gnb, const unsigned char* mark) {
  Call c(fn);
  ListArgs L = stage_lists(c, ix, iy, iz, nlo, nhi);
  if (L.ok) {
    const int r = order / 2;
    Box valid = isect(isect(gg, shrink(gphi, r, r, r)), gnb);
    Field<R> fx = c.out(gx, gg), fy = c.out(gy, gg), fz = c.out(gz, gg);
    Field<const R> fp = c.in(phi, gphi);
    if (order == 2)
      run_entries(c, "k_entries<central2>", L, nlo, nhi, marks(c, nb, gnb, mark), valid,
                  OpCentral2<R>{fx, fy, fz, fp, R(0.5 / (double)dx), R(0.5 / (double)dy), R(0.5 / (double)dz)});
    else
      run_entries(c, "k_entries<central4>", L, nlo, nhi, marks(c, nb, gnb, mark), valid,
                  OpCentral4<R>{fx, fy, fz, fp, R(0.0833333333333333333333 / (double)dx),
                                R(0.0833333333333333333333 / (double)dy), R(0.0833333333333333333333 / (double)dz)});
  }
  finish_local(c);
}

void lsm3dcentralgradorder2local_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* dx, const R* dy,
                                  const R* dz, const int* ix, const int* iy, const int* iz, const int* nlo,
                                  const int* nhi, const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  central_local("LSM3D_CENTRAL_GRAD_ORDER2_LOCAL", 2, gx, gy, gz, BOX(g), phi, BOX(phi), *dx, *dy, *dz, ix, iy, iz,
                *nlo, *nhi, nb, BOX(nb), mark_fb);
}

void lsm3dcentralgradorder4local_(R* gx, R* gy, R* gz, BOXP(g), const R* phi, BOXP(phi), const R* dx, const R* dy,
                                  const R* dz, const int* ix, const int* iy, const int* iz, const int* nlo,
                                  const int* nhi, const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  central_local("LSM3D_CENTRAL_GRAD_ORDER4_LOCAL", 4, gx, gy, gz, BOX(g), phi, BOX(phi), *dx, *dy, *dz, ix, iy, iz,
                *nlo, *nhi, nb, BOX(nb), mark_fb);
}

}  // extern "C"

namespace {

template <int ORDER, bool GAUSS>
void curvature_local(const char* fn, R* kappa, Box gk, const R* phi, Box gp, const R* px, const R* py, const R* pz,
                     const R* gm, Box gg, R dx, R dy, R dz, const int* ix, const int* iy, const int* iz, int nlo,
                     int nhi, const unsigned char* nb, Box gnb, const unsigned char* mark_fb) {
  Call c(fn);
  ListArgs L = stage_lists(c, ix, iy, iz, nlo, nhi);
  if (L.ok) {
    constexpr int r = ORDER / 2, rg = ORDER == 4 ? 2 : 0;
    OpCurvature<R, ORDER, GAUSS> op{c.out(kappa, gk), c.in(phi, gp), c.in(px, gg), c.in(py, gg),
                                    c.in(pz, gg),     c.in(gm, gg),  curv_coef<ORDER>(dx, dy, dz)};
    run_entries(c, "k_entries<curvature>", L, nlo, nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(gk, shrink(gp, r, r, r)), shrink(gg, rg, rg, rg)), gnb), op);
  }
  finish_local(c);
}

}  // namespace

extern "C" {

#define CURV_LOCAL_ARGS                                                                                       \
  R *kappa, BOXP(k), const R *phi, BOXP(phi), const R *px, const R *py, const R *pz, const R *gm, BOXP(g),    \
      const R *dx, const R *dy, const R *dz, const int *ix, const int *iy, const int *iz, const int *nlo,     \
      const int *nhi, const unsigned char *nb, BOXP(nb), const unsigned char *mark_fb
#define CURV_LOCAL_PASS                                                                                       \
  kappa, BOX(k), phi, BOX(phi), px, py, pz, gm, BOX(g), *dx, *dy, *dz, ix, iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb

void lsm3dcomputemeancurvatureorder2local_(CURV_LOCAL_ARGS) {
  curvature_local<2, false>("LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL", CURV_LOCAL_PASS);
}
void lsm3dcomputegaussiancurvatureorder2local_(CURV_LOCAL_ARGS) {
  curvature_local<2, true>("LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2_LOCAL", CURV_LOCAL_PASS);
}
// The reference's loop body of this routine divides the Gaussian numerator by
// |grad phi|^4 (lsm_curvature3d_local.f.in:403-417): it returns the
// fourth-order Gaussian curvature, and so does this drop-in.
void lsm3dcomputemeancurvatureorder4local_(CURV_LOCAL_ARGS) {
  curvature_local<4, true>("LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL", CURV_LOCAL_PASS);
}
void lsm3dcomputegaussiancurvatureorder4local_(CURV_LOCAL_ARGS) {
  curvature_local<4, true>("LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4_LOCAL", CURV_LOCAL_PASS);
}

void lsm3dlaplacianorder2local_(R* lap, BOXP(l), const R* phi, BOXP(phi), const R* dx, const R* dy, const R* dz,
                                const int* ix, const int* iy, const int* iz, const int* nlo, const int* nhi,
                                const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_LAPLACIAN_ORDER2_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gl = BOX(l), gp = BOX(phi), gnb = BOX(nb);
    const double x = *dx, y = *dy, z = *dz;
    OpLaplacian<R> op{c.out(lap, gl), c.in(phi, gp), R(1.0 / x / x), R(1.0 / y / y), R(1.0 / z / z)};
    run_entries(c, "k_entries<laplacian>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(gl, shrink(gp, 1, 1, 1)), gnb), op);
  }
  finish_local(c);
}

void lsm3dcomputeavegradphilocal_(R* ave, const R* phi, BOXP(phi), const R* dx, const R* dy, const R* dz,
                                  const int* ix, const int* iy, const int* iz, const int* nlo, const int* nhi,
                                  const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  double out[2] = {0.0, 0.0};
  if (L.ok) {
    const Box gp = BOX(phi), gnb = BOX(nb);
    RedAveGrad<R> op{c.in(phi, gp), R(0.5 / (double)*dx), R(0.5 / (double)*dy), R(0.5 / (double)*dz)};
    run_entries_reduce(c, "k_entries_reduce<avegrad>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                       isect(shrink(gp, 1, 1, 1), gnb), op, RED_LIST_SUM, out);
  }
  finish_local(c);
  if (!lsm::error_pending()) *ave = (out[1] > 0.0) ? R(out[0] / out[1]) : R(out[0]);
}

// The reference header declares the output grad_phi_mag const
// (lsm_spatial_derivatives3d_local.h:442); the prototype keeps that, and the
// routine writes it as the Fortran does.
void lsm3dgradientmagnitudelocal_(const R* gx, const R* gy, const R* gz, const R* mag_out, BOXP(g), const int* ix,
                                  const int* iy, const int* iz, const int* nlo, const int* nhi,
                                  const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  R* mag = const_cast<R*>(mag_out);
  Call c("LSM3D_GRADIENT_MAGNITUDE_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box g = BOX(g), gnb = BOX(nb);
    OpGradMag<R> op{c.out(mag, g), c.in(gx, g), c.in(gy, g), c.in(gz, g)};
    run_entries(c, "k_entries<gradmag>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb), isect(g, gnb), op);
  }
  finish_local(c);
}

// ---- TVD Runge-Kutta: lsm_tvd_runge_kutta3d_local.h ----------------------------
void lsm3drk1steplocal_(R* un, BOXP(un), const R* uc, BOXP(uc), const R* rhs, BOXP(r), const R* dt, const int* ix,
                        const int* iy, const int* iz, const int* nlo, const int* nhi, const unsigned char* nb,
                        BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_RK1_STEP_LOCAL", RK_EULER, un, BOX(un), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), *dt, ix, iy,
           iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}
void lsm3dtvdrk2stage1local_(R* un, BOXP(un), const R* uc, BOXP(uc), const R* rhs, BOXP(r), const R* dt,
                             const int* ix, const int* iy, const int* iz, const int* nlo, const int* nhi,
                             const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_TVD_RK2_STAGE1_LOCAL", RK_EULER, un, BOX(un), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), *dt, ix,
           iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}
void lsm3dtvdrk2stage2local_(R* un, BOXP(un), const R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs, BOXP(r),
                             const R* dt, const int* ix, const int* iy, const int* iz, const int* nlo,
                             const int* nhi, const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_TVD_RK2_STAGE2_LOCAL", RK2_STAGE2, un, BOX(un), u1, BOX(u1), uc, BOX(uc), rhs, BOX(r), *dt, ix,
           iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}
void lsm3dtvdrk3stage1local_(R* un, BOXP(un), const R* uc, BOXP(uc), const R* rhs, BOXP(r), const R* dt,
                             const int* ix, const int* iy, const int* iz, const int* nlo, const int* nhi,
                             const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_TVD_RK3_STAGE1_LOCAL", RK_EULER, un, BOX(un), nullptr, Box{}, uc, BOX(uc), rhs, BOX(r), *dt, ix,
           iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}
void lsm3dtvdrk3stage2local_(R* un, BOXP(un), const R* u1, BOXP(u1), const R* uc, BOXP(uc), const R* rhs, BOXP(r),
                             const R* dt, const int* ix, const int* iy, const int* iz, const int* nlo,
                             const int* nhi, const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_TVD_RK3_STAGE2_LOCAL", RK3_STAGE2, un, BOX(un), u1, BOX(u1), uc, BOX(uc), rhs, BOX(r), *dt, ix,
           iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}
void lsm3dtvdrk3stage3local_(R* un, BOXP(un), const R* u2, BOXP(u2), const R* uc, BOXP(uc), const R* rhs, BOXP(r),
                             const R* dt, const int* ix, const int* iy, const int* iz, const int* nlo,
                             const int* nhi, const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  rk_local("LSM3D_TVD_RK3_STAGE3_LOCAL", RK3_STAGE3, un, BOX(un), u2, BOX(u2), uc, BOX(uc), rhs, BOX(r), *dt, ix,
           iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}

// ---- level set equation RHS: lsm_level_set_evolution3d_local.h ------------------
void lsm3dzerooutlevelseteqnrhslocal_(R* rhs, BOXP(r), const int* ix, const int* iy, const int* iz,
                                      const int* nlo, const int* nhi) {
  Call c("LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box g = BOX(r);
    run_entries(c, "k_entries<zero>", L, *nlo, *nhi, no_marks(), g, OpZero<R>{c.out(rhs, g)});
  }
  finish_local(c);
}

void lsm3daddadvectiontermtolserhslocal_(R* rhs, BOXP(r), const R* px, const R* py, const R* pz, BOXP(g),
                                         const R* vx, const R* vy, const R* vz, BOXP(v), const int* ix,
                                         const int* iy, const int* iz, const int* nlo, const int* nhi,
                                         const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), gg = BOX(g), gv = BOX(v), gnb = BOX(nb);
    OpAdvection<R> op{c.out(rhs, gr), c.in(px, gg), c.in(py, gg), c.in(pz, gg), c.in(vx, gv), c.in(vy, gv),
                      c.in(vz, gv)};
    run_entries(c, "k_entries<advection>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(gr, gg), gv), gnb), op);
  }
  finish_local(c);
}

static void normal_vel_local(const char* fn, R* rhs, Box gr, const R* xp, const R* yp, const R* zp, Box gp,
                             const R* xm, const R* ym, const R* zm, Box gm, const R* vn, Box gv, R vn_const,
                             const int* ix, const int* iy, const int* iz, int nlo, int nhi,
                             const unsigned char* nb, Box gnb, const unsigned char* mark) {
  Call c(fn);
  ListArgs L = stage_lists(c, ix, iy, iz, nlo, nhi);
  if (L.ok) {
    Box valid = isect(isect(isect(gr, gp), gm), gnb);
    if (vn) valid = isect(valid, gv);
    OpNormalVel<R> op{c.out(rhs, gr), c.in(xp, gp), c.in(yp, gp), c.in(zp, gp), c.in(xm, gm),
                      c.in(ym, gm), c.in(zm, gm), vn ? c.in(vn, gv) : none(), vn_const};
    run_entries(c, "k_entries<normalvel>", L, nlo, nhi, marks(c, nb, gnb, mark), valid, op);
  }
  finish_local(c);
}

void lsm3daddnormalveltermtolserhslocal_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                         const R* xm, const R* ym, const R* zm, BOXP(gm), const R* vn, BOXP(v),
                                         const int* ix, const int* iy, const int* iz, const int* nlo,
                                         const int* nhi, const unsigned char* nb, BOXP(nb),
                                         const unsigned char* mark_fb) {
  normal_vel_local("LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL", rhs, BOX(r), xp, yp, zp, BOX(gp), xm, ym, zm,
                   BOX(gm), vn, BOX(v), R(0), ix, iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}

void lsm3daddconstnormalveltermtolserhslocal_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                              const R* xm, const R* ym, const R* zm, BOXP(gm), const R* vn,
                                              const int* ix, const int* iy, const int* iz, const int* nlo,
                                              const int* nhi, const unsigned char* nb, BOXP(nb),
                                              const unsigned char* mark_fb) {
  normal_vel_local("LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL", rhs, BOX(r), xp, yp, zp, BOX(gp), xm, ym,
                   zm, BOX(gm), nullptr, Box{}, *vn, ix, iy, iz, *nlo, *nhi, nb, BOX(nb), mark_fb);
}

void lsm3daddconstcurvtermtolserhslocal_(R* rhs, BOXP(r), const R* px, const R* py, const R* pz, BOXP(g1),
                                         const R* pxx, const R* pxy, const R* pxz, const R* pyy, const R* pyz,
                                         const R* pzz, BOXP(g2), const R* b, const int* ix, const int* iy,
                                         const int* iz, const int* nlo, const int* nhi, const unsigned char* nb,
                                         BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), g1 = BOX(g1), g2 = BOX(g2), gnb = BOX(nb);
    OpConstCurv<R> op{c.out(rhs, gr), c.in(px, g1),  c.in(py, g1),  c.in(pz, g1),  c.in(pxx, g2),
                      c.in(pxy, g2),  c.in(pxz, g2), c.in(pyy, g2), c.in(pyz, g2), c.in(pzz, g2), *b};
    run_entries(c, "k_entries<curv>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(gr, g1), g2), gnb), op);
  }
  finish_local(c);
}

// The reference header declares lse_rhs (in/out) const and kappa (in)
// non-const (lsm_level_set_evolution3d_local.h:375, 382); the prototype keeps
// those qualifiers, and the routine updates lse_rhs as the Fortran does.
void lsm3daddconstprecomputedcurvtermtolserhslocal_(const R* rhs_io, BOXP(r), R* kappa, BOXP(k), const R* gmag,
                                                    BOXP(g), const R* b, const int* ix, const int* iy,
                                                    const int* iz, const int* nlo, const int* nhi,
                                                    const unsigned char* nb, BOXP(nb),
                                                    const unsigned char* mark_fb) {
  R* rhs = const_cast<R*>(rhs_io);
  Call c("LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), gk = BOX(k), gg = BOX(g), gnb = BOX(nb);
    OpPrecomputedCurv<R> op{c.out(rhs, gr), c.in(kappa, gk), c.in(gmag, gg), *b};
    run_entries(c, "k_entries<precurv>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(gr, gk), gg), gnb), op);
  }
  finish_local(c);
}

void lsm3daddexternalandnormalveltermtolserhslocal_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp,
                                                    BOXP(gp), const R* xm, const R* ym, const R* zm, BOXP(gm),
                                                    const R* vn, const R* vx, const R* vy, const R* vz, BOXP(v),
                                                    const int* ix, const int* iy, const int* iz, const int* nlo,
                                                    const int* nhi, const unsigned char* nb, BOXP(nb),
                                                    const unsigned char* mark_fb) {
  Call c("LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), gpp = BOX(gp), gmm = BOX(gm), gv = BOX(v), gnb = BOX(nb);
    OpExtNormal<R> op{c.out(rhs, gr), c.in(xp, gpp), c.in(yp, gpp), c.in(zp, gpp), c.in(xm, gmm), c.in(ym, gmm),
                      c.in(zm, gmm),  c.in(vn, gv),  c.in(vx, gv),  c.in(vy, gv),  c.in(vz, gv)};
    run_entries(c, "k_entries<extnormal>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(isect(gr, gpp), gmm), gv), gnb), op);
  }
  finish_local(c);
}

// ---- reinitialization: lsm_reinitialization3d_local.h ---------------------------
void lsm3dcomputereinitializationeqnrhslocal_(R* rhs, BOXP(r), const R* phi, BOXP(phi), const R* phi0, BOXP(p0),
                                              const R* xp, const R* yp, const R* zp, BOXP(gp), const R* xm,
                                              const R* ym, const R* zm, BOXP(gm), const R* dx, const R* dy,
                                              const R* dz, const int* use_phi0, const int* ix, const int* iy,
                                              const int* iz, const int* nlo, const int* nhi,
                                              const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), gphi = BOX(phi), g0 = BOX(p0), gpp = BOX(gp), gmm = BOX(gm), gnb = BOX(nb);
    const bool u0 = (*use_phi0 == 1);
    const R m = max3(*dx, *dy, *dz);
    OpReinit<R> op{c.out(rhs, gr), c.in(phi, gphi), u0 ? c.in(phi0, g0) : none(), c.in(xp, gpp), c.in(yp, gpp),
                   c.in(zp, gpp),  c.in(xm, gmm),   c.in(ym, gmm),  c.in(zm, gmm), m * m, u0 ? 1 : 0};
    Box valid = isect(isect(isect(isect(gr, gphi), gpp), gmm), gnb);
    if (u0) valid = isect(valid, g0);
    run_entries(c, "k_entries<reinit>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb), valid, op);
  }
  finish_local(c);
}

void lsm3dcomputeorthogonalizationeqnrhslocal_(R* rhs, BOXP(r), const R* xp, const R* yp, const R* zp, BOXP(gp),
                                               const R* xm, const R* ym, const R* zm, BOXP(gm), const R* psi,
                                               BOXP(psi), const R* sxp, const R* syp, const R* szp, BOXP(sp),
                                               const R* sxm, const R* sym, const R* szm, BOXP(sm), const R* dx,
                                               const R* dy, const R* dz, const int* ix, const int* iy,
                                               const int* iz, const int* nlo, const int* nhi,
                                               const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  if (L.ok) {
    const Box gr = BOX(r), gpp = BOX(gp), gmm = BOX(gm), gps = BOX(psi), gsp = BOX(sp), gsm = BOX(sm),
              gnb = BOX(nb);
    const R m = max3(*dx, *dy, *dz);
    OpOrtho<R> op{c.out(rhs, gr),  c.in(xp, gpp),  c.in(yp, gpp),  c.in(zp, gpp),  c.in(xm, gmm),
                  c.in(ym, gmm),   c.in(zm, gmm),  c.in(psi, gps), c.in(sxp, gsp), c.in(syp, gsp),
                  c.in(szp, gsp),  c.in(sxm, gsm), c.in(sym, gsm), c.in(szm, gsm), m * m};
    run_entries(c, "k_entries<ortho>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                isect(isect(isect(isect(isect(isect(gr, gpp), gmm), gps), gsp), gsm), gnb), op);
  }
  finish_local(c);
}

// ---- reductions: lsm_utilities3d_local.h ---------------------------------------------
void lsm3dmaxnormdifflocal_(R* out, const R* f1, BOXP(f1), const R* f2, BOXP(f2), const int* ix, const int* iy,
                            const int* iz, const int* nlo, const int* nhi, const unsigned char* nb, BOXP(nb),
                            const unsigned char* mark_fb) {
  Call c("LSM3D_MAX_NORM_DIFF_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  double r[2] = {0.0, 0.0};
  if (L.ok) {
    const Box g1 = BOX(f1), g2 = BOX(f2), gnb = BOX(nb);
    RedMaxDiff<R> op{c.in(f1, g1), c.in(f2, g2)};
    run_entries_reduce(c, "k_entries_reduce<maxdiff>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                       isect(isect(g1, g2), gnb), op, RED_LIST_MAX, r);
  }
  finish_local(c);
  if (!lsm::error_pending()) *out = (r[0] > 0.0) ? R(r[0]) : R(0);  // the reference starts from 0
}

static void normal_dt_local(const char* fn, R* dt, const R* vn, Box gv, R vn_const, const R* xp, const R* yp,
                            const R* zp, Box gp, const R* xm, const R* ym, const R* zm, Box gm, const R* cv,
                            Box gcv, int cv_sgn, R dx, R dy, R dz, R cfl, const int* ix, const int* iy,
                            const int* iz, int nlo, int nhi, const unsigned char* nb, Box gnb,
                            const unsigned char* mark) {
  Call c(fn);
  ListArgs L = stage_lists(c, ix, iy, iz, nlo, nhi);
  double r[2] = {-1.0e308, 0.0};
  if (L.ok) {
    Box valid = isect(isect(gp, gm), gnb);
    if (vn) valid = isect(valid, gv);
    if (cv) valid = isect(valid, gcv);
    const R mdx = max3(dx, dy, dz);
    RedNormalDt<R> op{c.in(xp, gp), c.in(yp, gp), c.in(zp, gp), c.in(xm, gm), c.in(ym, gm), c.in(zm, gm),
                      vn ? c.in(vn, gv) : none(), cv ? c.in(cv, gcv) : none(), R(fabs(vn_const)), mdx * mdx,
                      R(1.0 / (double)dx), R(1.0 / (double)dy), R(1.0 / (double)dz), cv_sgn};
    run_entries_reduce(c, "k_entries_reduce<normaldt>", L, nlo, nhi, marks(c, nb, gnb, mark), valid, op,
                       RED_LIST_MAX, r);
  }
  finish_local(c);
  R mx = R(-1.0);
  if (r[0] > (double)mx) mx = R(r[0]);
  if (!lsm::error_pending()) *dt = cfl / (mx + Tol<R>::tiny);
}

void lsm3dcomputestableadvectiondtlocal_(R* dt, const R* vx, const R* vy, const R* vz, BOXP(v), const R* dx,
                                         const R* dy, const R* dz, const R* cfl, const int* ix, const int* iy,
                                         const int* iz, const int* nlo, const int* nhi, const unsigned char* nb,
                                         BOXP(nb), const unsigned char* mark_fb) {
  Call c("LSM3D_COMPUTE_STABLE_ADVECTION_DT_LOCAL");
  ListArgs L = stage_lists(c, ix, iy, iz, *nlo, *nhi);
  double r[2] = {-1.0e308, 0.0};
  if (L.ok) {
    const Box gv = BOX(v), gnb = BOX(nb);
    RedAdvDt<R> op{c.in(vx, gv), c.in(vy, gv), c.in(vz, gv), R(1.0 / (double)*dx), R(1.0 / (double)*dy),
                   R(1.0 / (double)*dz)};
    run_entries_reduce(c, "k_entries_reduce<advdt>", L, *nlo, *nhi, marks(c, nb, gnb, mark_fb),
                       isect(gv, gnb), op, RED_LIST_MAX, r);
  }
  finish_local(c);
  R mx = R(-1.0);
  if (r[0] > (double)mx) mx = R(r[0]);
  if (!lsm::error_pending()) *dt = *cfl / (mx + Tol<R>::tiny);
}

void lsm3dcomputestablenormalveldtlocal_(R* dt, const R* vn, BOXP(v), const R* xp, const R* yp, const R* zp,
                                         BOXP(gp), const R* xm, const R* ym, const R* zm, BOXP(gm), const R* dx,
                                         const R* dy, const R* dz, const R* cfl, const int* ix, const int* iy,
                                         const int* iz, const int* nlo, const int* nhi, const unsigned char* nb,
                                         BOXP(nb), const unsigned char* mark_fb) {
  normal_dt_local("LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL", dt, vn, BOX(v), R(0), xp, yp, zp, BOX(gp), xm, ym,
                  zm, BOX(gm), nullptr, Box{}, 0, *dx, *dy, *dz, *cfl, ix, iy, iz, *nlo, *nhi, nb, BOX(nb),
                  mark_fb);
}

void lsm3dcomputestablenormalveldtcontrolvolumelocal_(R* dt, const R* vn, BOXP(v), const R* xp, const R* yp,
                                                      const R* zp, BOXP(gp), const R* xm, const R* ym, const R* zm,
                                                      BOXP(gm), const R* cv, BOXP(cv), const int* cv_sgn,
                                                      const R* dx, const R* dy, const R* dz, const R* cfl,
                                                      const int* ix, const int* iy, const int* iz, const int* nlo,
                                                      const int* nhi, const unsigned char* nb, BOXP(nb),
                                                      const unsigned char* mark_fb) {
  normal_dt_local("LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME_LOCAL", dt, vn, BOX(v), R(0), xp, yp, zp,
                  BOX(gp), xm, ym, zm, BOX(gm), cv, BOX(cv), *cv_sgn, *dx, *dy, *dz, *cfl, ix, iy, iz, *nlo, *nhi,
                  nb, BOX(nb), mark_fb);
}

void lsm3dcomputestableconstnormalveldtlocal_(R* dt, const R* vn, const R* xp, const R* yp, const R* zp, BOXP(gp),
                                              const R* xm, const R* ym, const R* zm, BOXP(gm), const R* dx,
                                              const R* dy, const R* dz, const R* cfl, const int* ix, const int* iy,
                                              const int* iz, const int* nlo, const int* nhi,
                                              const unsigned char* nb, BOXP(nb), const unsigned char* mark_fb) {
  normal_dt_local("LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL", dt, nullptr, Box{}, *vn, xp, yp, zp, BOX(gp),
                  xm, ym, zm, BOX(gm), nullptr, Box{}, 0, *dx, *dy, *dz, *cfl, ix, iy, iz, *nlo, *nhi, nb, BOX(nb),
                  mark_fb);
}

}  // extern "C"
