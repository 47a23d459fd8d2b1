// Common device/host definitions for the gfx950 level-set kernels.
//
// Data layout (the reference's, lsm_spatial_derivatives3d.h:68-80): each field
// is a column-major array over its own ghost box a(ilo:ihi, jlo:jhi, klo:khi),
// i fastest.  Kernels see a field as a base pointer plus the ghost-box lower
// corner and 64-bit plane/row strides, so uneven ghost boxes and arbitrary
// signed lower bounds cost nothing.  Work always runs over a fill (or
// include) box that the host has checked lies inside every ghost box with the
// stencil's halo.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef LSM_REAL
#define LSM_REAL double
#endif
typedef LSM_REAL real_t;

namespace lsm {

struct Box {
  int ilo, ihi, jlo, jhi, klo, khi;
  __host__ __device__ int nx() const { return ihi - ilo + 1; }
  __host__ __device__ int ny() const { return jhi - jlo + 1; }
  __host__ __device__ int nz() const { return khi - klo + 1; }
  __host__ __device__ int64_t cells() const {
    return (nx() > 0 && ny() > 0 && nz() > 0) ? (int64_t)nx() * ny() * nz() : 0;
  }
};

// A field on its ghost box.  at(i,j,k) takes global (caller) indices.
template <typename T>
struct Field {
  T* p;
  int ilo, jlo, klo;
  int64_t sy, sz;  // row and plane strides in elements
  __device__ __forceinline__ int64_t off(int i, int j, int k) const {
    return (int64_t)(i - ilo) + sy * (int64_t)(j - jlo) + sz * (int64_t)(k - klo);
  }
  __device__ __forceinline__ T& at(int i, int j, int k) const { return p[off(i, j, k)]; }
};

template <typename T>
__host__ __device__ inline Field<T> make_field(T* p, const Box& gb) {
  Field<T> f;
  f.p = p;
  f.ilo = gb.ilo;
  f.jlo = gb.jlo;
  f.klo = gb.klo;
  f.sy = (int64_t)gb.nx();
  f.sz = (int64_t)gb.nx() * (int64_t)gb.ny();
  return f;
}

// Precision-dependent tolerances of the reference build (configure.ac:264-272):
// LSMLIB_ZERO_TOL and tiny_nonzero_number.
template <typename T> struct Tol;
template <> struct Tol<double> {
  static constexpr double zero = 1e-11;
  static constexpr double tiny = 1e-99;
};
template <> struct Tol<float> {
  static constexpr float zero = 1e-5f;
  static constexpr float tiny = 1e-35f;
};

// Launch geometry for the simple per-cell kernels: 64 x 4 threads over (i, j),
// k from the grid's z dimension (chunked when the box is deeper than 65535).
constexpr int kBX = 64, kBY = 4;

inline dim3 cell_grid(const Box& fb) {
  unsigned gz = (unsigned)fb.nz();
  if (gz > 65535u) gz = 65535u;
  return dim3((unsigned)((fb.nx() + kBX - 1) / kBX), (unsigned)((fb.ny() + kBY - 1) / kBY), gz);
}

#define LSM_CELL_LOOP_BEGIN(fb)                                        \
  const int i = (fb).ilo + (int)(blockIdx.x * blockDim.x + threadIdx.x); \
  const int j = (fb).jlo + (int)(blockIdx.y * blockDim.y + threadIdx.y); \
  if (i > (fb).ihi || j > (fb).jhi) return;                             \
  for (int k = (fb).klo + (int)blockIdx.z; k <= (fb).khi; k += (int)gridDim.z) {
#define LSM_CELL_LOOP_END }

}  // namespace lsm
