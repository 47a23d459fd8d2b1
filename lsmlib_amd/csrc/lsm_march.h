// Device building blocks of the fused stage kernels: XCD-aware tile mapping
// and the 2.5D z-march that keeps a (TY+6) x (TX+6) plane of the stage field
// in LDS and the seven z-neighbours of each column in registers.
#pragma once

#include "lsm_common.h"
#include "lsm_stencil.h"

namespace lsm {

constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, each with its own L2

// The hardware hands consecutive workgroup ids to the 8 XCDs round-robin.
// Remap so that XCD x runs a contiguous range of logical tiles (a slab of the
// grid), keeping the halo re-reads of neighbouring tiles in that XCD's L2.
// A bijection on [0, n) for any n.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n / kNumXcd, r = n % kNumXcd;
  const int x = b % kNumXcd, idx = b / kNumXcd;
  return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Tile geometry of one launch: fill box split into TX x TY columns and
// chunks of kz planes.
struct TileGrid {
  int ntx, nty, nkz;  // tiles along x, y and number of z chunks
  int kz;             // planes per chunk
};

inline TileGrid make_tile_grid(const Box& fb, int TX, int TY, int kz) {
  TileGrid g;
  g.ntx = (fb.nx() + TX - 1) / TX;
  g.nty = (fb.ny() + TY - 1) / TY;
  g.kz = kz;
  g.nkz = (fb.nz() + kz - 1) / kz;
  return g;
}

// Selects the five divided differences WENO5 needs on the upwind side and
// evaluates it once (no divergence across a velocity sign change).
// p[0..6] = phi(c-3..c+3).  vel > 0 takes the backward stencil
// (lsm_spatial_derivatives3d.f.in:2759-2766), otherwise the forward one
// (:2768-2775); |vel| < zero_tol gives 0 (:2756-2757).
template <typename T>
__device__ __forceinline__ T upwind_weno5_select(const T* p, T vel, T inv) {
  T d[6];
#pragma unroll
  for (int q = 0; q < 6; q++) d[q] = (p[q + 1] - p[q]) * inv;
  const bool bwd = vel > T(0);
  const T v1 = bwd ? d[0] : d[5];
  const T v2 = bwd ? d[1] : d[4];
  const T v3 = bwd ? d[2] : d[3];
  const T v4 = bwd ? d[3] : d[2];
  const T v5 = bwd ? d[4] : d[1];
  const T r = weno5_ref<T>(v1, v2, v3, v4, v5);
  return (fabs(vel) < Tol<T>::zero) ? T(0) : r;
}

// Writes `val` at (i,j,k) and, for each periodic axis whose image of the
// cell lies in the ghost layer, at the periodic images (up to 7 more cells).
template <typename T>
__device__ __forceinline__ void store_with_periodic_images(const Field<T>& out, const Box& fb, int i,
                                                           int j, int k, T val, int periodic, int g) {
  out.at(i, j, k) = val;
  if (!periodic) return;
  int xi[2] = {i, i}, yj[2] = {j, j}, zk[2] = {k, k};
  int nxs = 1, nys = 1, nzs = 1;
  if (periodic & 1) {
    if (i - fb.ilo < g) xi[nxs++] = i + fb.nx();
    else if (fb.ihi - i < g) xi[nxs++] = i - fb.nx();
  }
  if (periodic & 2) {
    if (j - fb.jlo < g) yj[nys++] = j + fb.ny();
    else if (fb.jhi - j < g) yj[nys++] = j - fb.ny();
  }
  if (periodic & 4) {
    if (k - fb.klo < g) zk[nzs++] = k + fb.nz();
    else if (fb.khi - k < g) zk[nzs++] = k - fb.nz();
  }
  for (int c = 0; c < nzs; c++)
    for (int b = 0; b < nys; b++)
      for (int a = 0; a < nxs; a++)
        if (a | b | c) out.at(xi[a], yj[b], zk[c]) = val;
}

}  // namespace lsm
