// One-sided Hamilton-Jacobi ENO1/2/3 and WENO5 derivatives along one axis.
//
// Every routine takes the seven samples p[0..6] = phi(c-3 .. c+3) along the
// axis through cell c (only p[3-R .. 3+R] are read, R = scheme radius) and
// the reciprocal spacing inv = 1/dx computed on the host as the reference
// does (inv_dx = 1.0d0/dx, lsm_spatial_derivatives3d.f.in:276).
//
// The undivided differences are formed exactly as lsm3dComputeDn builds them
// (lsm_spatial_derivatives3d.f.in:111-121): with d1[q] = D1(c+q-2),
//   D1(i) = phi(i) - phi(i-1)           (face i-1/2 stored at i)
//   D2(i) = D1(i+1) - D1(i)             (cell centred)
//   D3(i) = D2(i) - D2(i-1)             (face centred)
// Those are pure subtractions, so the ENO stencil choices (strict '<' on
// |D2|, |D3|) are bit-identical to the reference's whatever the compiler
// fuses elsewhere.
#pragma once

#include "lsm_common.h"

namespace lsm {

enum Scheme { ENO1 = 1, ENO2 = 2, ENO3 = 3, WENO5 = 5 };

template <int S> struct SchemeRadius { static constexpr int R = (S == ENO1) ? 1 : (S == ENO2) ? 2 : 3; };

// WENO5 combination of five divided differences, in the reference's order
// (lsm_spatial_derivatives3d.f.in:1292-1326).
template <typename T>
__device__ __forceinline__ T weno5_ref(T v1, T v2, T v3, T v4, T v5) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  const T one_third = T(1.0 / 3.0), seven_sixths = T(7.0 / 6.0), eleven_sixths = T(11.0 / 6.0);
  const T one_sixth = T(1.0 / 6.0), five_sixths = T(5.0 / 6.0);
  const T thirteen_twelfths = T(13.0 / 12.0), one_fourth = T(0.25);
  T m = v1 * v1;
  m = fmax(m, v2 * v2);
  m = fmax(m, v3 * v3);
  m = fmax(m, v4 * v4);
  m = fmax(m, v5 * v5);
  const T eps = T(1e-6) * m + Tol<T>::tiny;
  const T p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  const T p2 = -one_sixth * v2 + five_sixths * v3 + one_third * v4;
  const T p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  const T t1 = v1 - T(2) * v2 + v3, t2 = v1 - T(4) * v2 + T(3) * v3;
  const T t3 = v2 - T(2) * v3 + v4, t4 = v2 - v4;
  const T t5 = v3 - T(2) * v4 + v5, t6 = T(3) * v3 - T(4) * v4 + v5;
  const T S1 = thirteen_twelfths * (t1 * t1) + one_fourth * (t2 * t2);
  const T S2 = thirteen_twelfths * (t3 * t3) + one_fourth * (t4 * t4);
  const T S3 = thirteen_twelfths * (t5 * t5) + one_fourth * (t6 * t6);
  const T s1 = S1 + eps, s2 = S2 + eps, s3 = S3 + eps;
  T a1 = T(0.1) / (s1 * s1);
  T a2 = T(0.6) / (s2 * s2);
  T a3 = T(0.3) / (s3 * s3);
  const T inv_sum_a = T(1) / (a1 + a2 + a3);
  a1 = a1 * inv_sum_a;
  a2 = a2 * inv_sum_a;
  a3 = a3 * inv_sum_a;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

// Reciprocal used by the fast WENO5 form: the hardware v_rcp_f32 (about
// 1 ulp) in FP32, an exact division in FP64.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }

// The same WENO5 combination as weno5_ref, rearranged to about a third of
// the operations and two reciprocals, on UNDIVIDED differences D1..D5
// (= v1..v5 times dx); the result is the undivided derivative (multiply by
// 1/dx).  With the second differences a = D1-D2, b = D2-D3, c = D3-D4,
// d = D4-D5, the Jiang-Shu indicators times 12 are quadratic forms
//   12 S1 = 13(a-b)^2 + 3(a-3b)^2 = 16 (a (a - 2.75 b) + 2.5 b^2)
//   12 S2 = 13(b-c)^2 + 3(b+c)^2  = 16 (b (b - 1.25 c) + c^2)
//   12 S3 = 13(c-d)^2 + 3(3c-d)^2 = 16 (d (d - 2.75 c) + 2.5 c^2)
// and eps = 1e-6 max(v_k^2) + tiny becomes E = 12e-6 max|D|^2 + tiny_dx2
// (tiny_dx2 = 12 * tiny * dx^2 from the caller).  The weights
// alpha_k = d_k/(S_k + eps)^2, d = (0.1, 0.6, 0.3), are formed on
// sigma_k = (12 S_k + E)/E >= 1, so that with P_kl = sigma_k^2 sigma_l^2
//   w1 = P23 / Den,  w3 = 3 P12 / Den,  Den = P23 + 6 P13 + 3 P12
// (sigma_k <= ~4e7 keeps Den finite in FP32).  The result is
//   phi_x = p2 + w1 (p1 - p2) + w3 (p3 - p2),
//   p2 = (-D2 + 5 D3 + 2 D4)/6 = D3 - (b + 2c)/6,
//   p1 - p2 = (a - 2b + c)/3,  p3 - p2 = (b - 2c + d)/6.
// The multiply-adds are written as explicit fma so that every compiled copy
// (unrolled planes, tile shapes) rounds alike.
// max(|a|, |b|, |c|, |d|, |e|) of finite values in two v_max3_f32 with
// absolute-value source modifiers.  The compiler's fmax first quiets
// signalling NaNs of operands that come from another basic block or a loop
// phi (one v_max per operand); these operands are differences of finite
// samples, so the quieting is not needed.
__device__ __forceinline__ float max_abs5(float a, float b, float c, float d, float e) {
  float t, m;
  asm("v_max3_f32 %0, |%1|, |%2|, |%3|" : "=v"(t) : "v"(a), "v"(b), "v"(c));
  asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(m) : "v"(t), "v"(d), "v"(e));
  return m;
}
__device__ __forceinline__ double max_abs5(double a, double b, double c, double d, double e) {
  return fmax(fmax(fmax(fabs(a), fabs(b)), fabs(c)), fmax(fabs(d), fabs(e)));
}

template <typename T>
__device__ __forceinline__ T weno5_fast(T D1, T D2, T D3, T D4, T D5, T tiny_dx2) {
#pragma clang fp contract(off)
  const T a = D1 - D2, b = D2 - D3, c = D3 - D4, d = D4 - D5;
  const T m = max_abs5(D1, D2, D3, D4, D5);
  const T ie = fast_rcp(fma(T(0.75e-6), m * m, tiny_dx2 * T(1.0 / 16.0)));  // 16 / E
  const T bb = b * b, cc = c * c;
  const T q1 = fma(a, fma(T(-2.75), b, a), T(2.5) * bb);
  const T q2 = fma(b, fma(T(-1.25), c, b), cc);
  const T q3 = fma(d, fma(T(-2.75), c, d), T(2.5) * cc);
  const T s1 = fma(q1, ie, T(1)), s2 = fma(q2, ie, T(1)), s3 = fma(q3, ie, T(1));
  const T t1 = s1 * s1, t2 = s2 * s2, t3 = s3 * s3;
  const T p23 = t2 * t3, p13 = t1 * t3, p12 = t1 * t2;
  const T rden = fast_rcp(fma(T(3), p12, fma(T(6), p13, p23)));
  const T x = fma(p23 * (fma(T(-2), b, a) + c), T(1.0 / 3.0), (p12 * (fma(T(-2), c, b) + d)) * T(0.5));
  return fma(x, rden, fma(fma(T(2), c, b), T(-1.0 / 6.0), D3));
}

// Plus and minus WENO5 (weno5_fast) of one axis from the six first
// differences d[0..5] around the cell, sharing what the two stencils have in
// common: with e_q = d[q+1] - d[q] the minus stencil's second differences
// (a, b, c, d) are (-e0, -e1, -e2, -e3) and the plus stencil's
// (e4, e3, e2, e1); c^2 = e2^2 and 2.5 c^2 are the same in both, and max|D|
// shares max(|d1|..|d4|).  Negation and the maximum are exact, so both
// results equal weno5_fast(d0..d4) and weno5_fast(d5..d1) bit for bit.
// Returns the two undivided derivatives (multiply by 1/dx).  tf is the
// weight offset tiny * dx^2 / 16 (the caller forms it once).
template <typename T>
__device__ __forceinline__ void weno5_fast_pm(const T* d, T tf, T& plus, T& minus) {
#pragma clang fp contract(off)
  const T e0 = d[1] - d[0], e1 = d[2] - d[1], e2 = d[3] - d[2], e3 = d[4] - d[3], e4 = d[5] - d[4];
  T mid, mm, mp;
  if constexpr (sizeof(T) == 4) {
    asm("v_max3_f32 %0, |%1|, |%2|, |%3|" : "=v"(mid) : "v"(d[1]), "v"(d[2]), "v"(d[3]));
    asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(mm) : "v"(mid), "v"(d[0]), "v"(d[4]));
    asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(mp) : "v"(mid), "v"(d[4]), "v"(d[5]));
  } else {
    mid = fmax(fmax(fabs(d[1]), fabs(d[2])), fabs(d[3]));
    mm = fmax(fmax(mid, fabs(d[0])), fabs(d[4]));
    mp = fmax(fmax(mid, fabs(d[4])), fabs(d[5]));
  }
  const T iem = fast_rcp(fma(T(0.75e-6), mm * mm, tf)), iep = fast_rcp(fma(T(0.75e-6), mp * mp, tf));
  const T cc = e2 * e2, cc25 = T(2.5) * cc;
  // minus: (a, b, c, d) = -(e0, e1, e2, e3)
  {
    const T q1 = fma(e0, fma(T(-2.75), e1, e0), T(2.5) * (e1 * e1));
    const T q2 = fma(e1, fma(T(-1.25), e2, e1), cc);
    const T q3 = fma(e3, fma(T(-2.75), e2, e3), cc25);
    const T s1 = fma(q1, iem, T(1)), s2 = fma(q2, iem, T(1)), s3 = fma(q3, iem, T(1));
    const T t1 = s1 * s1, t2 = s2 * s2, t3 = s3 * s3;
    const T p23 = t2 * t3, p13 = t1 * t3, p12 = t1 * t2;
    const T rden = fast_rcp(fma(T(3), p12, fma(T(6), p13, p23)));
    // a - 2b + c = -(e0 - 2 e1 + e2), b - 2c + d = -(e1 - 2 e2 + e3), b + 2c = -(e1 + 2 e2)
    const T x = fma(p23 * -(fma(T(-2), e1, e0) + e2), T(1.0 / 3.0), (p12 * -(fma(T(-2), e2, e1) + e3)) * T(0.5));
    minus = fma(x, rden, fma(-fma(T(2), e2, e1), T(-1.0 / 6.0), d[2]));
  }
  // plus: (a, b, c, d) = (e4, e3, e2, e1)
  {
    const T q1 = fma(e4, fma(T(-2.75), e3, e4), T(2.5) * (e3 * e3));
    const T q2 = fma(e3, fma(T(-1.25), e2, e3), cc);
    const T q3 = fma(e1, fma(T(-2.75), e2, e1), cc25);
    const T s1 = fma(q1, iep, T(1)), s2 = fma(q2, iep, T(1)), s3 = fma(q3, iep, T(1));
    const T t1 = s1 * s1, t2 = s2 * s2, t3 = s3 * s3;
    const T p23 = t2 * t3, p13 = t1 * t3, p12 = t1 * t2;
    const T rden = fast_rcp(fma(T(3), p12, fma(T(6), p13, p23)));
    const T x = fma(p23 * (fma(T(-2), e3, e4) + e2), T(1.0 / 3.0), (p12 * (fma(T(-2), e2, e3) + e1)) * T(0.5));
    plus = fma(x, rden, fma(fma(T(2), e2, e3), T(-1.0 / 6.0), d[3]));
  }
}

// Godunov |grad phi|*^2 of the reinitialization equation, as godunov_star_sq:
// per axis max(-p, m, 0) (s > 0) or max(p, -m, 0) (otherwise), which is the
// reference's max(max(-p, 0), max(m, 0)) exactly (a maximum rounds nothing),
// in one three-operand maximum per axis.
template <typename T>
__device__ __forceinline__ T godunov_star_sq_max3(T s, T xp, T yp, T zp, T xm, T ym, T zm) {
#pragma clang fp contract(off)
  const T z = T(0);
  const bool pos = s > z;
  const T gx = pos ? fmax(fmax(-xp, xm), z) : fmax(fmax(xp, -xm), z);
  const T gy = pos ? fmax(fmax(-yp, ym), z) : fmax(fmax(yp, -ym), z);
  const T gz = pos ? fmax(fmax(-zp, zm), z) : fmax(fmax(zp, -zm), z);
  return gx * gx + gy * gy + gz * gz;
}

// The same from undivided one-sided derivatives, each axis's |.|* scaled by its
// 1/dx after the maxima: multiplying by a positive factor is monotonic and odd
// under rounding, so max(-a f, b f, 0) = f max(-a, b, 0) exactly and the result
// equals godunov_star_sq_max3 on the scaled derivatives bit for bit.
template <typename T>
__device__ __forceinline__ T godunov_star_sq_undivided(T s, T xp, T yp, T zp, T xm, T ym, T zm, T ix, T iy, T iz) {
#pragma clang fp contract(off)
  const T z = T(0);
  const bool pos = s > z;
  const T gx = ix * (pos ? fmax(fmax(-xp, xm), z) : fmax(fmax(xp, -xm), z));
  const T gy = iy * (pos ? fmax(fmax(-yp, ym), z) : fmax(fmax(yp, -ym), z));
  const T gz = iz * (pos ? fmax(fmax(-zp, zm), z) : fmax(fmax(zp, -zm), z));
  return gx * gx + gy * gy + gz * gz;
}

// d1[q] = D1(c+q-2), q = 0..5 (only those the scheme needs are formed).
template <typename T, int S>
__device__ __forceinline__ void first_differences(const T* p, T* d1) {
  if (S == ENO1) {
    d1[2] = p[3] - p[2];
    d1[3] = p[4] - p[3];
  } else if (S == ENO2) {
#pragma unroll
    for (int q = 1; q <= 4; q++) d1[q] = p[q + 1] - p[q];
  } else {
#pragma unroll
    for (int q = 0; q <= 5; q++) d1[q] = p[q + 1] - p[q];
  }
}

// Plus (forward) derivative at c, lsm_spatial_derivatives3d.f.in:299-308 (ENO1),
// :518-527 (ENO2), :838-868 (ENO3), :1286-1326 (WENO5).
template <typename T, int S>
__device__ __forceinline__ T hj_plus(const T* p, T inv) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  T d1[6];
  first_differences<T, S>(p, d1);
  if (S == ENO1) {
    return d1[3] * inv;
  } else if (S == ENO2) {
    const T d2c = d1[3] - d1[2], d2p = d1[4] - d1[3];
    return (fabs(d2c) < fabs(d2p)) ? (d1[3] - T(0.5) * d2c) * inv : (d1[3] - T(0.5) * d2p) * inv;
  } else if (S == ENO3) {
    const T d2m = d1[2] - d1[1], d2c = d1[3] - d1[2], d2p = d1[4] - d1[3], d2pp = d1[5] - d1[4];
    const T d3c = d2c - d2m, d3p = d2p - d2c, d3pp = d2pp - d2p;
    T v = d1[3];
    if (fabs(d2c) < fabs(d2p)) {
      v = v - T(0.5) * d2c;
      v = (fabs(d3c) < fabs(d3p)) ? v - T(1.0 / 6.0) * d3c : v - T(1.0 / 6.0) * d3p;
    } else {
      v = v - T(0.5) * d2p;
      v = (fabs(d3p) < fabs(d3pp)) ? v + T(1.0 / 3.0) * d3p : v + T(1.0 / 3.0) * d3pp;
    }
    return v * inv;
  } else {
    return weno5_ref<T>(d1[5] * inv, d1[4] * inv, d1[3] * inv, d1[2] * inv, d1[1] * inv);
  }
}

// Minus (backward) derivative at c, lsm_spatial_derivatives3d.f.in:299-308 (ENO1),
// :529-535 (ENO2), :872-896 (ENO3), :1330-1370 (WENO5).
template <typename T, int S>
__device__ __forceinline__ T hj_minus(const T* p, T inv) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  T d1[6];
  first_differences<T, S>(p, d1);
  if (S == ENO1) {
    return d1[2] * inv;
  } else if (S == ENO2) {
    const T d2m = d1[2] - d1[1], d2c = d1[3] - d1[2];
    return (fabs(d2m) < fabs(d2c)) ? (d1[2] + T(0.5) * d2m) * inv : (d1[2] + T(0.5) * d2c) * inv;
  } else if (S == ENO3) {
    const T d2mm = d1[1] - d1[0], d2m = d1[2] - d1[1], d2c = d1[3] - d1[2], d2p = d1[4] - d1[3];
    const T d3m = d2m - d2mm, d3c = d2c - d2m, d3p = d2p - d2c;
    T v = d1[2];
    if (fabs(d2m) < fabs(d2c)) {
      v = v + T(0.5) * d2m;
      v = (fabs(d3m) < fabs(d3c)) ? v + T(1.0 / 3.0) * d3m : v + T(1.0 / 3.0) * d3c;
    } else {
      v = v + T(0.5) * d2c;
      v = (fabs(d3c) < fabs(d3p)) ? v - T(1.0 / 6.0) * d3c : v - T(1.0 / 6.0) * d3p;
    }
    return v * inv;
  } else {
    return weno5_ref<T>(d1[0] * inv, d1[1] * inv, d1[2] * inv, d1[3] * inv, d1[4] * inv);
  }
}

// Upwind selection (lsm_spatial_derivatives3d.f.in:1735-1743, 2756-2775):
// |vel| < zero_tol -> 0; vel > 0 -> backward; otherwise forward.
template <typename T, int S>
__device__ __forceinline__ T hj_upwind(const T* p, T vel, T inv) {
  if (fabs(vel) < Tol<T>::zero) return T(0);
  return (vel > T(0)) ? hj_minus<T, S>(p, inv) : hj_plus<T, S>(p, inv);
}

// Godunov |grad phi|* of the reinitialization equation for the sign of s
// (lsm_reinitialization3d.f.in:168-185); returns |grad phi|*^2
template <typename T>
__device__ __forceinline__ T godunov_star_sq(T s, T xp, T yp, T zp, T xm, T ym, T zm) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  const T z = T(0);
  T gx, gy, gz;
  if (s > z) {
    gx = fmax(fmax(-xp, z), fmax(xm, z));
    gy = fmax(fmax(-yp, z), fmax(ym, z));
    gz = fmax(fmax(-zp, z), fmax(zm, z));
  } else {
    gx = fmax(fmax(xp, z), fmax(-xm, z));
    gy = fmax(fmax(yp, z), fmax(-ym, z));
    gz = fmax(fmax(zp, z), fmax(-zm, z));
  }
  return gx * gx + gy * gy + gz * gz;
}

// Godunov |grad phi|^2 for normal speed vn, lsm_level_set_evolution3d.f.in:206-225
template <typename T>
__device__ __forceinline__ T godunov_norm_sq(T vn, T xp, T yp, T zp, T xm, T ym, T zm) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  const T z = T(0);
  if (vn > z) {
    const T a = fmax(xm, z), b = fmin(xp, z), c = fmax(ym, z), d = fmin(yp, z), e = fmax(zm, z),
            f = fmin(zp, z);
    return fmax(a * a, b * b) + fmax(c * c, d * d) + fmax(e * e, f * f);
  }
  const T a = fmin(xm, z), b = fmax(xp, z), c = fmin(ym, z), d = fmax(yp, z), e = fmin(zm, z),
          f = fmax(zp, z);
  return fmax(a * a, b * b) + fmax(c * c, d * d) + fmax(e * e, f * f);
}

// kappa |grad phi| from central first and second derivatives,
// lsm_level_set_evolution3d.f.in:433-447 (0 where |grad phi|^2 < zero_tol),
// every product and sum rounded as the reference rounds it.
template <typename T>
__device__ __forceinline__ T curvature_term(T x, T y, T z, T xx, T xy, T xz, T yy, T yz, T zz, T zero_tol) {
#pragma clang fp contract(off)
  const T g2 = (x * x + y * y) + z * z;
  if (g2 < zero_tol) return T(0);
  T n = (xx * y) * y + (yy * x) * x;
  n = n - ((T(2) * xy) * x) * y;
  n = n + (xx * z) * z;
  n = n + (zz * x) * x;
  n = n - ((T(2) * xz) * x) * z;
  n = n + (yy * z) * z;
  n = n + (zz * y) * y;
  n = n - ((T(2) * yz) * y) * z;
  return n / g2;
}

// Mean and Gaussian curvature from phi, its given first derivatives and
// |grad phi| (lsm_curvature3d.f.in).  Coefficients, computed on the host as
// FP64 quotients rounded to LSMLIB_REAL as the reference's assignments do:
//   order 2: f2 = 1/dx_d^2, fm = 0.25/(dx dy), 0.25/(dx dz), 0.25/(dy dz)
//   order 4: f2 = (1/12)/dx_d^2, fm = (1/12)/dx, (1/12)/dy, (1/12)/dz
template <typename T>
struct CurvCoef {
  T f2[3];
  T fm[3];
};

template <int ORDER, typename T>
inline CurvCoef<T> curv_coef(T dx, T dy, T dz) {
  const double c12 = 0.0833333333333333333333;  // the reference's 1/12 literal
  const double d[3] = {double(dx), double(dy), double(dz)};
  CurvCoef<T> c;
  for (int q = 0; q < 3; q++) c.f2[q] = T((ORDER == 2 ? 1.0 : c12) / d[q] / d[q]);
  if (ORDER == 2) {
    c.fm[0] = T(0.25 / d[0] / d[1]);
    c.fm[1] = T(0.25 / d[0] / d[2]);
    c.fm[2] = T(0.25 / d[1] / d[2]);
  } else {
    for (int q = 0; q < 3; q++) c.fm[q] = T(c12 / d[q]);
  }
  return c;
}

// Second derivatives at (i,j,k).  Order 2 (:90-101): the pure ones are FP64
// sums (the 2.0d0 literal promotes them) rounded to LSMLIB_REAL, the mixed ones
// LSMLIB_REAL.  Order 4 (:330-349): LSMLIB_REAL throughout; the mixed ones
// differentiate phi_x along y, phi_y along z and phi_z along x.
template <typename T, int ORDER>
__device__ __forceinline__ void curv_hessian(const Field<const T>& phi, const Field<const T>& px,
                                             const Field<const T>& py, const Field<const T>& pz, int i, int j,
                                             int k, const CurvCoef<T>& c, T& xx, T& yy, T& zz, T& xy, T& xz,
                                             T& yz) {
#pragma clang fp contract(off)
  if (ORDER == 2) {
    const double p0 = 2.0 * double(phi.at(i, j, k));
    xx = T(((double(phi.at(i + 1, j, k)) - p0) + double(phi.at(i - 1, j, k))) * double(c.f2[0]));
    yy = T(((double(phi.at(i, j + 1, k)) - p0) + double(phi.at(i, j - 1, k))) * double(c.f2[1]));
    zz = T(((double(phi.at(i, j, k + 1)) - p0) + double(phi.at(i, j, k - 1))) * double(c.f2[2]));
    xy = (((phi.at(i + 1, j + 1, k) - phi.at(i + 1, j - 1, k)) - phi.at(i - 1, j + 1, k)) +
          phi.at(i - 1, j - 1, k)) * c.fm[0];
    xz = (((phi.at(i + 1, j, k + 1) - phi.at(i + 1, j, k - 1)) - phi.at(i - 1, j, k + 1)) +
          phi.at(i - 1, j, k - 1)) * c.fm[1];
    yz = (((phi.at(i, j + 1, k + 1) - phi.at(i, j + 1, k - 1)) - phi.at(i, j - 1, k + 1)) +
          phi.at(i, j - 1, k - 1)) * c.fm[2];
  } else {
    const T s16 = T(16), s30 = T(30), e8 = T(8), c30 = s30 * phi.at(i, j, k);
    xx = ((((-phi.at(i + 2, j, k) + s16 * phi.at(i + 1, j, k)) - c30) - phi.at(i - 2, j, k)) +
          s16 * phi.at(i - 1, j, k)) * c.f2[0];
    yy = ((((-phi.at(i, j + 2, k) + s16 * phi.at(i, j + 1, k)) - c30) - phi.at(i, j - 2, k)) +
          s16 * phi.at(i, j - 1, k)) * c.f2[1];
    zz = ((((-phi.at(i, j, k + 2) + s16 * phi.at(i, j, k + 1)) - c30) - phi.at(i, j, k - 2)) +
          s16 * phi.at(i, j, k - 1)) * c.f2[2];
    xy = (((-px.at(i, j + 2, k) + e8 * px.at(i, j + 1, k)) + px.at(i, j - 2, k)) - e8 * px.at(i, j - 1, k)) *
         c.fm[1];
    yz = (((-py.at(i, j, k + 2) + e8 * py.at(i, j, k + 1)) + py.at(i, j, k - 2)) - e8 * py.at(i, j, k - 1)) *
         c.fm[2];
    xz = (((-pz.at(i + 2, j, k) + e8 * pz.at(i + 1, j, k)) + pz.at(i - 2, j, k)) - e8 * pz.at(i - 1, j, k)) *
         c.fm[0];
  }
}

// kappa at (i,j,k): the mean curvature's numerator over |grad phi|^3
// (:103-120) or the Gaussian's over |grad phi|^4 (:224-239); 0 where that
// power is below zero_tol.
template <typename T, int ORDER, bool GAUSS>
__device__ __forceinline__ T curvature_at(const Field<const T>& phi, const Field<const T>& px,
                                          const Field<const T>& py, const Field<const T>& pz,
                                          const Field<const T>& gm, int i, int j, int k, const CurvCoef<T>& c) {
#pragma clang fp contract(off)
  const T g = gm.at(i, j, k);
  const T den = GAUSS ? ((g * g) * g) * g : (g * g) * g;
  if (den < Tol<T>::zero) return T(0);
  T xx, yy, zz, xy, xz, yz;
  curv_hessian<T, ORDER>(phi, px, py, pz, i, j, k, c, xx, yy, zz, xy, xz, yz);
  const T x = px.at(i, j, k), y = py.at(i, j, k), z = pz.at(i, j, k);
  T n;
  if (GAUSS) {
    n = (x * x) * (yy * zz - yz * yz) + (y * y) * (xx * zz - xz * xz);
    n = n + (z * z) * (xx * yy - xy * xy);
    T m = (x * y) * (xz * yz - xy * zz) + (y * z) * (xy * xz - yz * xx);
    m = m + (x * z) * (xy * yz - xz * yy);
    n = n + T(2) * m;
  } else {
    n = (xx * y) * y + (yy * x) * x;
    n = n - ((T(2) * xy) * x) * y;
    n = n + (xx * z) * z;
    n = n + (zz * x) * x;
    n = n - ((T(2) * xz) * x) * z;
    n = n + (yy * z) * z;
    n = n + (zz * y) * y;
    n = n - ((T(2) * yz) * y) * z;
  }
  return n / den;
}

// CFL speed of a normal velocity, lsm_utilities3d.f.in:404-429
template <typename T>
__device__ __forceinline__ T h_over_dx(T avn, T xp, T yp, T zp, T xm, T ym, T zm, T mdx2, T ix, T iy,
                                       T iz) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  const T px = fmax(fabs(xp), fabs(xm)), py = fmax(fabs(yp), fabs(ym)), pz = fmax(fabs(zp), fabs(zm));
  const T ng = sqrt(px * px + py * py + pz * pz + mdx2);
  return avn / ng * (px * ix + py * iy + pz * iz);
}

// Godunov selection for the combined external + normal velocity,
// lsm_level_set_evolution3d.f.in:630-650
template <typename T>
__device__ __forceinline__ T ext_normal_select(T vc, T vn, T dp, T dm) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  const T hp = vc + vn * dp, hm = vc + vn * dm;
  if (hp > T(0)) return (hm > T(0)) ? dm : -(vc / vn);
  if (hm < T(0)) return dp;
  return (fabs(hp) > fabs(hm)) ? dp : dm;
}

// Lax-Friedrichs numerical Hamiltonians (no counterpart in the reference,
// whose level-set RHS routines are upwind or Godunov only).  From the plus
// and minus one-sided derivatives p, m of one axis and a dissipation
// coefficient alpha >= |dH/dphi_axis|:
//   advection, H = V . grad(phi), per axis
//     H_a = (0.5 (v - alpha)) p + (0.5 (v + alpha)) m
//   = v (p + m)/2 - alpha (p - m)/2.  With alpha = |v| one of the two
//   coefficients is exactly 0 and the other exactly v (2v and 0.5 * 2v are
//   exact), so H_a = v m (v > 0) or v p (v < 0) bit for bit: the upwind term
//   of LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS with the upwind selection of
//   lsm_spatial_derivatives3d.f.in:2756-2810 (where |v| >= zero_tol).  In
//   general H_a = upwind + 0.5 (|v| - alpha) (p - m).
//   normal speed, H = vn |grad phi|:
//     H = vn |(p + m)/2| - sum_a alpha (p_a - m_a)/2,   alpha = |vn|
//   which on a planar signed distance (p = m, |grad phi| = 1) is vn.
template <typename T>
__host__ __device__ inline T lf_axis(T v, T alpha, T p, T m) {
#pragma clang fp contract(off)
  return (T(0.5) * (v - alpha)) * p + (T(0.5) * (v + alpha)) * m;
}

template <typename T>
__host__ __device__ inline T lf_normal(T vn, T alpha, T xp, T yp, T zp, T xm, T ym, T zm) {
#pragma clang fp contract(off)
  const T ax = T(0.5) * (xp + xm), ay = T(0.5) * (yp + ym), az = T(0.5) * (zp + zm);
  const T visc = (alpha * (T(0.5) * (xp - xm)) + alpha * (T(0.5) * (yp - ym))) + alpha * (T(0.5) * (zp - zm));
  return vn * sqrt((ax * ax + ay * ay) + az * az) - visc;
}

// TVD Runge-Kutta stage combinations (lsm_tvd_runge_kutta3d.f), given the
// current solution uc, the stage input us and the stage right-hand side r.
// UNBIASED (the FP32 fused stages) writes TVD-RK3 stage 3 as
// uc + 2/3 (us + dt r - uc): in FP32 the rounded 1/3 and 2/3 sum to
// 1 + 3e-8, so the reference's form scales phi by that factor every step
// (the reference float build does, lsm_tvd_runge_kutta3d.f:481-488); here the
// rounded 2/3 multiplies only the step's change.  The other stage weights are
// exact in binary.
enum RKForm { RK_EULER = 1, RK2_STAGE2 = 22, RK3_STAGE2 = 32, RK3_STAGE3 = 33 };

// The FP32 fused stages write every form as one branch-free blend,
// uc + w (us + dt r - uc), with w = 1 (forward Euler / stage 1, where us =
// uc), 1/2 (TVD-RK2 stage 2), 1/4 (TVD-RK3 stage 2) or 2/3 (TVD-RK3 stage 3,
// the unbiased form above); the same values as rk_combine up to FP32
// rounding.  The fused multiply-adds are explicit so every compiled copy
// rounds alike.
template <typename T>
__host__ __device__ inline T rk_blend_weight(int form) {
  switch (form) {
    case RK2_STAGE2:
      return T(0.5);
    case RK3_STAGE2:
      return T(0.25);
    case RK3_STAGE3:
      return T(2.0 / 3.0);
    default:
      return T(1);
  }
}

template <typename T>
__device__ __forceinline__ T rk_blend(T uc, T us, T r, T dt, T w) {
  return fma(w, fma(dt, r, us) - uc, uc);
}

template <typename T, bool UNBIASED = false>
__device__ __forceinline__ T rk_combine(int form, T uc, T us, T r, T dt) {
#pragma clang fp contract(off)  // products and sums rounded separately, as the reference's
  switch (form) {
    case RK2_STAGE2:  // :240-241
      return T(0.5) * (uc + us + dt * r);
    case RK3_STAGE2:  // :403-404
      return T(0.75) * uc + T(0.25) * (us + dt * r);
    case RK3_STAGE3:  // :486-488
      if (UNBIASED) return uc + T(2.0 / 3.0) * ((us + dt * r) - uc);
      return T(1.0 / 3.0) * uc + T(2.0 / 3.0) * (us + dt * r);
    default:  // u_next = u_cur + dt*rhs, :78
      return uc + dt * r;
  }
}

}  // namespace lsm
