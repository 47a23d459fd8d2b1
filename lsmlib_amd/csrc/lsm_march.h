// Device building blocks of the fused stage kernels: XCD-aware tile mapping,
// the boundary-aware loader and the 2.5D z-march that keeps a (TY+6) x (TX+6)
// plane of the stage field in LDS and the seven z-neighbours of each column
// in registers.
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

// Tile geometry: fill box split into TX x TY columns and chunks of kz planes.
struct TileGrid {
  int ntx, nty, nkz;  // tiles along x, y and number of z chunks
  int kz;             // planes per chunk
  __host__ __device__ int count() const { return ntx * nty * nkz; }
};

inline TileGrid make_tile_grid(const Box& fb, int TX, int TY, int kz) {
  TileGrid g;
  g.ntx = (fb.nx() + TX - 1) / TX;
  g.nty = (fb.ny() + TY - 1) / TY;
  g.kz = kz;
  g.nkz = (fb.nz() + kz - 1) / kz;
  return g;
}

// Ghost-cell codes per face (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi):
//   0 read the stored ghost cell, 1 copy (homogeneous Neumann), 2 linear,
//   3 signed linear extrapolation, 5 periodic.
// A cross-shaped stencil centred in the fill box leaves it through one face
// at a time, so every ghost value it reads is the single-face extrapolation
// of two interior cells (lsm_boundary_conditions3d.f:66-77, 213-225) or a
// periodic image; synthesising it here gives the same value the reference's
// fill-then-differentiate sequence reads, without a fill pass.
struct GhostRule {
  int code[6];
  int any;  // nonzero if some face is synthesised
};

// The product and sum are rounded separately, as in the reference (no fused
// multiply-add), so extrapolated ghost values are bit-identical to its fill.
template <typename T>
__device__ __forceinline__ T extrapolate(int code, T pb, T p1, int dist) {
#pragma clang fp contract(off)
  if (code == 1) return pb;
  T slope = (code == 2) ? pb - p1 : copysign(T(1), pb) * fabs(pb - p1);
  return pb + slope * T(dist);
}

// extrapolate() without branches, for codes that vary between lanes: code 0
// (a stored cell or a periodic image resolved to its source) and code 1
// return pb; codes 2 and 3 the same roundings as extrapolate() (the slope
// sign(pb) |pb - p1| is copysign(|pb - p1|, pb) exactly).  Selects instead of
// branches keep the loads of a plane free of control flow, so the next
// plane's loads stay in flight while a plane is computed.
template <typename T>
__device__ __forceinline__ T extrapolate_sel(int code, T pb, T p1, int dist) {
#pragma clang fp contract(off)
  const T d = pb - p1;
  const T slope = (code == 3) ? copysign(fabs(d), pb) : d;
  const T e = pb + slope * T(dist);
  return (code >= 2) ? e : pb;
}


// Writes `val` at (i,j,k) and, for each axis in `periodic` whose image of
// the cell lies in the g-wide ghost layer, at the periodic images (up to 7
// more cells).
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

// ---------------------------------------------------------------------------
// One fused Runge-Kutta stage over one tile.
// ---------------------------------------------------------------------------
enum StageKind { STAGE_ADVECT_FIELD = 0, STAGE_ADVECT_CONST = 1, STAGE_REINIT = 2, STAGE_ADVECT_LF = 3 };

// Stage kinds that read a stored velocity field.
template <int KIND>
constexpr bool reads_velocity() {
  return KIND == STAGE_ADVECT_FIELD || KIND == STAGE_ADVECT_LF;
}

template <typename T>
struct StageArgs {
  Field<T> out;
  Field<const T> ps;          // stage input
  Field<const T> pc;          // phi at the start of the step (p == nullptr: use ps)
  Field<const T> vx, vy, vz;  // velocity (advection, field form)
  Field<const T> phi0;        // reinit: frozen phi0 for the smoothed sign
  T cvx, cvy, cvz;            // constant velocity
  Box fb;
  TileGrid tg;
  GhostRule ghost;
  T ix, iy, iz, dt, dx_sq;
  T zero_tol;    // upwind switch / reinit cut-off (LSMLIB_ZERO_TOL of the build by default)
  T tdx2[3];     // 12 * tiny * dx^2 per axis (the WENO5 epsilon floor in undivided units)
  int form;      // RKForm
  T rkb;         // rk_blend_weight(form), used by the FP32 (FASTW) stages
  int periodic;  // bit mask: also write phi_out's periodic ghost images
  GhostRule oghost;  // extrapolation codes (1-3) of phi_out's face ghost layers to write (IMAGES)
  int use_phi0;  // reinit sign from phi0 (1) or phi_s (0)
  int alpha_mode;  // Lax-Friedrichs stages: 0 alpha = |v_a| per cell, 1 the global alpha[a]
  T alpha[3];
};

// ---------------------------------------------------------------------------
// Per-cell derivatives.  FASTW selects the rearranged WENO5 (weno5_fast)
// instead of the reference-order form.
// ---------------------------------------------------------------------------
// Upwind WENO5 derivative from the six first differences d[q] = p[q+1] - p[q]
// of the samples p[0..6] around the cell.
template <typename T, bool FASTW>
__device__ __forceinline__ T upwind_weno5_d(const T* d, T vel, T inv, T tiny_dx2, T zero_tol) {
  const bool bwd = vel > T(0);
  const T D1 = bwd ? d[0] : d[5], D2 = bwd ? d[1] : d[4], D3 = bwd ? d[2] : d[3];
  const T D4 = bwd ? d[3] : d[2], D5 = bwd ? d[4] : d[1];
  T r;
  if (FASTW) r = weno5_fast<T>(D1, D2, D3, D4, D5, tiny_dx2) * inv;
  else r = weno5_ref<T>(D1 * inv, D2 * inv, D3 * inv, D4 * inv, D5 * inv);
  return (fabs(vel) < zero_tol) ? T(0) : r;
}

// The same selection decided once per wavefront where it can be: when every
// active lane takes the backward stencil, or every lane the forward one, the
// WENO5 runs on that side without per-lane selects; when |vel| < zero_tol on
// every lane the derivative is 0 and no WENO5 is evaluated (the reference
// evaluates none there either, lsm_spatial_derivatives3d.f.in:2756-2758).
// Mixed wavefronts take upwind_weno5_d.  Every path computes the same value
// for a lane as upwind_weno5_d.
template <typename T, bool FASTW>
__device__ __forceinline__ T upwind_weno5_wave(const T* d, T vel, T inv, T tiny_dx2, T zero_tol) {
  const uint64_t live = __builtin_amdgcn_read_exec();
  const uint64_t nz = __builtin_amdgcn_ballot_w64(!(fabs(vel) < zero_tol));
  const uint64_t gt = __builtin_amdgcn_ballot_w64(vel > T(0));
  if (nz == 0) return T(0);
  if (nz == live) {
    if ((gt & live) == live) {
      if (FASTW) return weno5_fast<T>(d[0], d[1], d[2], d[3], d[4], tiny_dx2) * inv;
      return weno5_ref<T>(d[0] * inv, d[1] * inv, d[2] * inv, d[3] * inv, d[4] * inv);
    }
    if ((gt & live) == 0) {
      if (FASTW) return weno5_fast<T>(d[5], d[4], d[3], d[2], d[1], tiny_dx2) * inv;
      return weno5_ref<T>(d[5] * inv, d[4] * inv, d[3] * inv, d[2] * inv, d[1] * inv);
    }
  }
  return upwind_weno5_d<T, FASTW>(d, vel, inv, tiny_dx2, zero_tol);
}

template <typename T, bool FASTW>
__device__ __forceinline__ T upwind_weno5(const T* p, T vel, T inv, T tiny_dx2, T zero_tol) {
  T d[6];
#pragma unroll
  for (int q = 0; q < 6; q++) d[q] = p[q + 1] - p[q];
  return upwind_weno5_d<T, FASTW>(d, vel, inv, tiny_dx2, zero_tol);
}

// Plus and minus WENO5 derivatives from the six first differences
// (hj_plus / hj_minus for WENO5: lsm_spatial_derivatives3d.f.in:1286-1370).
template <typename T, bool FASTW>
__device__ __forceinline__ void hj_weno5_pm_d(const T* d, T inv, T tiny_dx2, T& plus, T& minus) {
  if (FASTW) {
    T p, m;
    weno5_fast_pm<T>(d, tiny_dx2 * T(1.0 / 16.0), p, m);
    plus = p * inv;
    minus = m * inv;
  } else {
    plus = weno5_ref<T>(d[5] * inv, d[4] * inv, d[3] * inv, d[2] * inv, d[1] * inv);
    minus = weno5_ref<T>(d[0] * inv, d[1] * inv, d[2] * inv, d[3] * inv, d[4] * inv);
  }
}

template <typename T, bool FASTW>
__device__ __forceinline__ void hj_weno5_pm(const T* p, T inv, T tiny_dx2, T& plus, T& minus) {
  T d[6];
#pragma unroll
  for (int q = 0; q < 6; q++) d[q] = p[q + 1] - p[q];
  hj_weno5_pm_d<T, FASTW>(d, inv, tiny_dx2, plus, minus);
}

// ---------------------------------------------------------------------------
// The z-march of one tile.  The workgroup (TX*TY threads, 1-D) owns tile
// (bx, by) x z-chunk bz: column (i, j) per thread, planes k0..k1 bottom to
// top; a (TY+6) x (TX+7) plane of the stage field sits in LDS and the z
// neighbours of each column in registers.
//
// Addressing is kept off the vector ALU: every field plane is a buffer
// resource built from a workgroup-uniform plane pointer (advanced by one
// plane per step in scalar registers) and read with a 32-bit per-lane byte
// offset fixed for the whole march.  Halo loads are unconditional (lanes
// without a halo cell read their own column), and periodic ghost images are
// written only by tiles and planes within 3 cells of a periodic face.
//
// SYNTH: halo values outside the fill box follow the ghost rule a.ghost per
// face (extrapolation or periodic image) instead of stored ghost cells.  A
// cross-shaped stencil leaves the fill box through one face at a time, so
// only three kinds of loads can need it: the x-halo column and y-halo rows
// of a lane (a per-lane LoadRule), and the z neighbours of the column near
// the z faces (uniform per plane).
// ---------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* p) {
  // raw buffer, no stride, no range limit, 32-bit data format (gfx9 word 3)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, 0x00020000);
}

typedef unsigned int lsm_u32x2 __attribute__((ext_vector_type(2)));

// Cache policy: kCached for the stage field (its halo is re-read by the
// neighbouring tiles), kStream (the non-temporal bit) for data touched once:
// velocity, phi_cur, phi0 and the output.
constexpr int kCached = 0, kStream = 2;

template <typename T, int POLICY = kCached>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, POLICY));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, POLICY));
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, 0, kStream);
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(lsm_u32x2, v), r, (int)off, 0, kStream);
}

// The three ghost cells beyond a face cell of the output: extrapolate(code,
// pb, p1, d) at distance d = 1..3 from the face value pb and its inward
// neighbour p1, the values the reference's face fill stores there
// (lsm_boundary_conditions3d.f:66-77, 213-225).  Ghost cell d lies in the
// plane at byte address plane + d * pstep, at byte offset off + d * ostep
// (x and y faces step the offset, z faces the plane), so the stores need
// no 64-bit per-lane address arithmetic.
template <typename T>
__device__ __forceinline__ void store_face_ghosts(const char* plane, int64_t pstep, uint32_t off, int32_t ostep,
                                                  int code, T pb, T p1) {
#pragma unroll
  for (int d = 1; d <= 3; d++)
    buf_store(extrapolate<T>(code, pb, p1, d), plane_rsrc(plane + d * pstep), off + (uint32_t)(d * ostep));
}

template <typename T>
__device__ __forceinline__ uint32_t plane_byte_offset(const Field<const T>& f, int i, int j) {
  return (uint32_t)((i - f.ilo) + (int)f.sy * (j - f.jlo)) * (uint32_t)sizeof(T);
}

// First byte of plane k of a field, and the byte distance between planes.
template <typename T>
__device__ __forceinline__ const char* plane_ptr(const Field<const T>& f, int k) {
  return reinterpret_cast<const char*>(f.p) + (int64_t)(k - f.klo) * f.sz * (int64_t)sizeof(T);
}

// How a lane obtains one halo value of the stage field in every plane: the
// stored cell at byte offset a (code 0), or the single-face extrapolation
// (codes 1-3) of the fill-box cells at a (on the face) and b (next to it) at
// distance dist.  Periodic faces are resolved to a wrapped offset here.
struct LoadRule {
  uint32_t a, b;
  int code, dist;
};

template <typename T, bool SYNTH>
__device__ __forceinline__ LoadRule load_rule(const Field<const T>& f, const Box& fb, const GhostRule& g, int ci,
                                              int cj) {
  if (SYNTH) {
    if (g.code[0] == 5 && ci < fb.ilo) ci += fb.nx();
    if (g.code[1] == 5 && ci > fb.ihi) ci -= fb.nx();
    if (g.code[2] == 5 && cj < fb.jlo) cj += fb.ny();
    if (g.code[3] == 5 && cj > fb.jhi) cj -= fb.ny();
    if (ci < fb.ilo && g.code[0])
      return {plane_byte_offset<T>(f, fb.ilo, cj), plane_byte_offset<T>(f, fb.ilo + 1, cj), g.code[0], fb.ilo - ci};
    if (ci > fb.ihi && g.code[1])
      return {plane_byte_offset<T>(f, fb.ihi, cj), plane_byte_offset<T>(f, fb.ihi - 1, cj), g.code[1], ci - fb.ihi};
    if (cj < fb.jlo && g.code[2])
      return {plane_byte_offset<T>(f, ci, fb.jlo), plane_byte_offset<T>(f, ci, fb.jlo + 1), g.code[2], fb.jlo - cj};
    if (cj > fb.jhi && g.code[3])
      return {plane_byte_offset<T>(f, ci, fb.jhi), plane_byte_offset<T>(f, ci, fb.jhi - 1), g.code[3], cj - fb.jhi};
  }
  const uint32_t o = plane_byte_offset<T>(f, ci, cj);
  return {o, o, 0, 0};
}

// A halo value loaded one plane ahead: both cells of the rule are fetched
// (b = a when no extrapolation applies) and combined only when consumed, so
// the loads stay in flight across the previous plane's arithmetic.
template <typename T, bool SYNTH>
struct Pending {
  T va, vb;
  __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, const LoadRule& lr) {
    va = buf_load<T>(r, lr.a);
    if (SYNTH) vb = buf_load<T>(r, lr.b);
  }
  __device__ __forceinline__ T value(const LoadRule& lr) const {
    if (SYNTH) return extrapolate_sel<T>(lr.code, va, vb, lr.dist);
    return va;
  }
};

// The planes a column value of plane kk is formed from (uniform): both kk
// inside the fill box, where the lane's own-column rule applies, or, past a
// synthesised z face, the face plane and the one next to it (or the periodic
// source plane) at the lane's own column with the face's code.
struct PlaneRule {
  const char* pa;
  const char* pb;
  int code, dist;
  bool face;
};

// IMAGES = false compiles out the periodic-image stores (callers whose
// a.periodic is always 0, such as the patch level); FACES = true compiles in
// the extrapolated face-ghost writes of a.oghost (stages that read stored
// ghosts and keep their output's face ghosts), which the other launches do
// not carry in their registers.
// The LDS plane is double-buffered: plane k is written to buffer k & 1 (of
// the chunk-relative index), so one barrier per plane suffices -- the reads
// of a buffer (plane k - 2) all precede the barrier of plane k - 1, which
// every wave passes before writing the buffer again at plane k.
template <typename T, int TX, int TY>
using StageTile = T[2][TY + 6][TX + 7];

template <typename T, int TX, int TY, int KIND, bool FASTW, bool SYNTH, bool IMAGES = true, bool FACES = false>
__device__ __forceinline__ void march_tile(const StageArgs<T>& a, int bx, int by, int bz,
                                           StageTile<T, TX, TY>& tiles) {
  // no implicit contraction: the FP32 arithmetic that fuses writes fma
  // explicitly, so unrolled copies of a plane round alike
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const Box& fb = a.fb;
  const int i0 = fb.ilo + bx * TX, j0 = fb.jlo + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = fb.klo + bz * a.tg.kz;
  const int k1 = min(k0 + a.tg.kz - 1, fb.khi);
  const bool active = (i <= fb.ihi) && (j <= fb.jhi);
  const int ic = min(i, fb.ihi + 3), jc = min(j, fb.jhi + 3);  // clamped (unused values)
  const GhostRule& g = a.ghost;
  const uint32_t oc = plane_byte_offset<T>(a.ps, ic, jc);
  // The column's own value in planes inside the fill box.  Lanes past the
  // last column or row of a partial tile hold the x- or y-halo of the active
  // cells in LDS, so under SYNTH their values follow the ghost rule too.
  const LoadRule rc = load_rule<T, SYNTH>(a.ps, fb, g, ic, jc);
  const bool do_x = tx < 6;
  const int xi = (tx < 3) ? i0 - 3 + tx : min(i0 + TX + tx - 3, fb.ihi + 3);
  const LoadRule rx = do_x ? load_rule<T, SYNTH>(a.ps, fb, g, xi, jc) : LoadRule{oc, oc, 0, 0};
  constexpr int NR = (6 + TY - 1) / TY;
  LoadRule ry[NR];
  int tyr[NR];
  bool doy[NR];
#pragma unroll
  for (int n = 0; n < NR; n++) {
    const int r = ty + n * TY;
    doy[n] = r < 6;
    const int yj = !doy[n] ? jc : (r < 3) ? j0 - 3 + r : min(j0 + TY + (r - 3), fb.jhi + 3);
    tyr[n] = !doy[n] ? 0 : (r < 3) ? r : TY + r;
    ry[n] = doy[n] ? load_rule<T, SYNTH>(a.ps, fb, g, ic, yj) : LoadRule{oc, oc, 0, 0};
  }
  const int txl = (tx < 3) ? tx : TX + tx;
  const uint32_t ov = reads_velocity<KIND>() ? plane_byte_offset<T>(a.vx, ic, jc) : 0u;
  const uint32_t opc = a.pc.p ? plane_byte_offset<T>(a.pc, ic, jc) : oc;
  const uint32_t op0 = (KIND == STAGE_REINIT && a.use_phi0) ? plane_byte_offset<T>(a.phi0, ic, jc) : oc;
  const uint32_t oo = (uint32_t)((ic - a.out.ilo) + (int)a.out.sy * (jc - a.out.jlo)) * (uint32_t)sizeof(T);
  const T tiny_dx2x = a.tdx2[0], tiny_dx2y = a.tdx2[1], tiny_dx2z = a.tdx2[2];
  // FP32 reinitialization: its loop-invariant factors (the WENO weight offsets
  // and 1/dx) copied to per-lane registers, since a VALU instruction with a
  // scalar-register operand issues at half the rate of one without
  T tfx = tiny_dx2x * T(1.0 / 16.0), tfy = tiny_dx2y * T(1.0 / 16.0), tfz = tiny_dx2z * T(1.0 / 16.0);
  T ivx = a.ix, ivy = a.iy, ivz = a.iz;
  if constexpr (FASTW && KIND == STAGE_REINIT)
    asm volatile("" : "+v"(tfx), "+v"(tfy), "+v"(tfz), "+v"(ivx), "+v"(ivy), "+v"(ivz));
  // periodic images: only tiles within 3 cells of a periodic x or y face
  const bool img_xy = IMAGES && (((a.periodic & 1) && (i0 - fb.ilo < 3 || fb.ihi - (i0 + TX - 1) < 3)) ||
                                 ((a.periodic & 2) && (j0 - fb.jlo < 3 || fb.jhi - (j0 + TY - 1) < 3)));
  // extrapolated face ghosts of phi_out: the tiles holding the face cells
  // write them from the face cell and its inward neighbour, which the host
  // guarantees lie in the same tile (x, y) and z-chunk (z)
  const int* oc6 = a.oghost.code;
  const bool f_xlo = FACES && a.oghost.any && oc6[0] && i0 == fb.ilo;
  const bool f_xhi = FACES && a.oghost.any && oc6[1] && i0 + TX - 1 >= fb.ihi;
  const bool f_ylo = FACES && a.oghost.any && oc6[2] && j0 == fb.jlo;
  const bool f_yhi = FACES && a.oghost.any && oc6[3] && j0 + TY - 1 >= fb.jhi;
  const bool f_xy = f_xlo || f_xhi || f_ylo || f_yhi;
  const bool f_z = FACES && a.oghost.any && ((oc6[4] && k0 == fb.klo) || (oc6[5] && k1 == fb.khi));
  T v_prev = T(0);  // phi_out of the column in the previous plane

  // plane pointers, advanced by one plane per step
  const int64_t s_ps = a.ps.sz * (int64_t)sizeof(T), s_out = a.out.sz * (int64_t)sizeof(T);
  const int64_t s_vel = a.vx.sz * (int64_t)sizeof(T), s_pc = a.pc.sz * (int64_t)sizeof(T);
  const int64_t s_p0 = a.phi0.sz * (int64_t)sizeof(T);
  const int64_t s_pc1 = a.pc.p ? s_pc : s_ps, s_p01 = (KIND == STAGE_REINIT && a.use_phi0) ? s_p0 : s_ps;
  const char* p_ps = plane_ptr<T>(a.ps, k0);
  const char* p_top = plane_ptr<T>(a.ps, k0 + 3);
  const char* p_vx = reads_velocity<KIND>() ? plane_ptr<T>(a.vx, k0) : nullptr;
  const char* p_vy = reads_velocity<KIND>() ? plane_ptr<T>(a.vy, k0) : nullptr;
  const char* p_vz = reads_velocity<KIND>() ? plane_ptr<T>(a.vz, k0) : nullptr;
  // absent optional fields load the column's value 3 planes up, the cell
  // the top load of the same plane reads (a first-level cache hit; the value
  // is unused), so that every load is issued unconditionally
  const bool has_pc = a.pc.p != nullptr, has_p0 = KIND == STAGE_REINIT && a.use_phi0;
  const char* p_pc = has_pc ? plane_ptr<T>(a.pc, k0) : p_top;
  const char* p_p0 = has_p0 ? plane_ptr<T>(a.phi0, k0) : p_top;
  char* p_out = reinterpret_cast<char*>(a.out.p) + (int64_t)(k0 - a.out.klo) * s_out;

  // z column: zv[q] = phi(k + q), q = 0..3, and zd[q] = phi(k + q - 2) -
  // phi(k + q - 3), q = 0..5 (one new value and difference per plane).  The
  // loads of plane k + 1 are issued before the arithmetic of plane k (one
  // plane of software pipelining).  FP32 planes are unrolled by two (the
  // column shifts and the pipelined values then need no register copies);
  // with no implicit contraction both copies compute identically.
  constexpr int U = FASTW ? 2 : 1;
  T zv[U + 3], zd[U + 5];
  // the column's value in plane kk, applying the z-face rule outside the fill
  // box (kk is uniform; there the value only feeds the column's own z
  // derivative, so lanes outside the fill box may read any stored cell)
  auto zrule = [&](int kk) -> PlaneRule {
    if (SYNTH && kk < fb.klo && g.code[4]) {
      if (g.code[4] == 5) {
        const char* p = plane_ptr<T>(a.ps, kk + fb.nz());
        return {p, p, 0, 0, true};
      }
      return {plane_ptr<T>(a.ps, fb.klo), plane_ptr<T>(a.ps, fb.klo + 1), g.code[4], fb.klo - kk, true};
    }
    if (SYNTH && kk > fb.khi && g.code[5]) {
      if (g.code[5] == 5) {
        const char* p = plane_ptr<T>(a.ps, kk - fb.nz());
        return {p, p, 0, 0, true};
      }
      return {plane_ptr<T>(a.ps, fb.khi), plane_ptr<T>(a.ps, fb.khi - 1), g.code[5], kk - fb.khi, true};
    }
    const char* p = plane_ptr<T>(a.ps, kk);
    return {p, p, 0, 0, false};
  };
  // the column's value of plane kk: both cells loaded, the rule applied when
  // the value is consumed (SYNTH); otherwise the stored cell
  struct ColValue {
    T va, vb;
    int code, dist;
  };
  auto col_load = [&](int kk, ColValue& c) {
    const PlaneRule z = zrule(kk);
    c.va = buf_load<T>(plane_rsrc(z.pa), z.face ? oc : rc.a);
    if (SYNTH) {
      c.vb = buf_load<T>(plane_rsrc(z.pb), z.face ? oc : rc.b);
      c.code = z.face ? z.code : rc.code;
      c.dist = z.face ? z.dist : rc.dist;
    }
  };
  auto col_value = [&](const ColValue& c) -> T {
    if (SYNTH) return extrapolate_sel<T>(c.code, c.va, c.vb, c.dist);
    return c.va;
  };
  {
    ColValue z0[6];
#pragma unroll
    for (int q = 0; q < 6; q++) col_load(k0 - 3 + q, z0[q]);
    T z1[6];
#pragma unroll
    for (int q = 0; q < 6; q++) z1[q] = col_value(z0[q]);
#pragma unroll
    for (int q = 0; q < 5; q++) zd[q] = z1[q + 1] - z1[q];
#pragma unroll
    for (int q = 0; q < 3; q++) zv[q] = z1[q + 3];
  }
  // plane values of the next plane to process (halo and top values pending)
  // U sets of pending values: plane k consumes set k % U (chunk-relative)
  // and fetches plane k + 1 into the other, so that with the planes
  // unrolled by two every set stays in its registers across the loop and
  // no copy of an in-flight load (which would wait for it) is needed
  Pending<T, SYNTH> phx[U], phy[U][NR];
  ColValue ptop[U];
  T ux[U], uy[U], uz[U], pcv[U], p0v[U];
#pragma unroll
  for (int q = 0; q < U; q++) ux[q] = uy[q] = uz[q] = pcv[q] = p0v[q] = T(0);
  // loads plane kn's values (and the column value 3 planes up) into set q
  auto fetch = [&](int kn, int q) {
    const __amdgpu_buffer_rsrc_t rps = plane_rsrc(p_ps);
    if (SYNTH) {
      col_load(kn + 3, ptop[q]);
    } else {
      ptop[q].va = buf_load<T>(plane_rsrc(p_top), oc);
    }
    phx[q].load(rps, rx);
#pragma unroll
    for (int n = 0; n < NR; n++) phy[q][n].load(rps, ry[n]);
    if (reads_velocity<KIND>()) {
      ux[q] = buf_load<T, kStream>(plane_rsrc(p_vx), ov);
      uy[q] = buf_load<T, kStream>(plane_rsrc(p_vy), ov);
      uz[q] = buf_load<T, kStream>(plane_rsrc(p_vz), ov);
    } else if (KIND == STAGE_ADVECT_CONST) {
      ux[q] = a.cvx;
      uy[q] = a.cvy;
      uz[q] = a.cvz;
    } else if (KIND == STAGE_REINIT) {
      p0v[q] = buf_load<T, kStream>(plane_rsrc(p_p0), op0);
    }
    pcv[q] = buf_load<T, kStream>(plane_rsrc(p_pc), opc);
  };
  auto advance = [&]() {
    p_ps += s_ps;
    p_top += s_ps;
    if (reads_velocity<KIND>()) {
      p_vx += s_vel;
      p_vy += s_vel;
      p_vz += s_vel;
    }
    p_pc += s_pc1;
    if (KIND == STAGE_REINIT) p_p0 += s_p01;
  };
  if (k0 <= k1) fetch(k0, 0);

  int par = 0;  // LDS buffer of the plane (U == 1; with U == 2 it is u)
  for (int kg = k0; kg <= k1; kg += U) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = kg + u;
      if (k > k1) break;  // uniform
      T (*tile)[TX + 7] = tiles[U == 2 ? u : par];
      par ^= 1;
      const int cs = u % U, ns = (u + 1) % U;  // consumed and refilled sets (static)
      zv[u + 3] = col_value(ptop[cs]);
      const T hx = phx[cs].value(rx);
      T hy[NR];
#pragma unroll
      for (int n = 0; n < NR; n++) hy[n] = phy[cs][n].value(ry[n]);
      tile[ty + 3][tx + 3] = zv[u];
      if (do_x) tile[ty + 3][txl] = hx;
#pragma unroll
      for (int n = 0; n < NR; n++)
        if (doy[n]) tile[tyr[n]][tx + 3] = hy[n];
      zd[u + 5] = zv[u + 3] - zv[u + 2];
      const T cux = ux[cs], cuy = uy[cs], cuz = uz[cs], cpc = pcv[cs], cp0 = p0v[cs];
      __syncthreads();
      // the next plane's loads; at the last plane the pointers stay and the
      // loads re-read plane k1 (unused), so the fetch is unconditional
      if (k < k1) advance();  // uniform
      fetch(k < k1 ? k + 1 : k1, ns);

      T v = T(0);
      if (active) {
        T dx_[6], dy_[6];
        {
          T px[7], py[7];
#pragma unroll
          for (int q = 0; q < 7; q++) {
            px[q] = tile[ty + 3][tx + q];
            py[q] = tile[ty + q][tx + 3];
          }
#pragma unroll
          for (int q = 0; q < 6; q++) {
            dx_[q] = px[q + 1] - px[q];
            dy_[q] = py[q + 1] - py[q];
          }
        }
        const T* dz_ = zd + u;
        const T us = zv[u];
        const T uc = a.pc.p ? cpc : us;
        T rhs;
        if (KIND == STAGE_REINIT) {
          T xp, xm, yp, ym, zp, zm;
          if constexpr (FASTW) {  // undivided: 1/dx is applied after the Godunov maxima
            weno5_fast_pm<T>(dx_, tfx, xp, xm);
            weno5_fast_pm<T>(dy_, tfy, yp, ym);
            weno5_fast_pm<T>(dz_, tfz, zp, zm);
          } else {
            hj_weno5_pm_d<T, FASTW>(dx_, a.ix, tiny_dx2x, xp, xm);
            hj_weno5_pm_d<T, FASTW>(dy_, a.iy, tiny_dx2y, yp, ym);
            hj_weno5_pm_d<T, FASTW>(dz_, a.iz, tiny_dx2z, zp, zm);
          }
          // lsm_reinitialization3d.f.in:150-268
          const T cur = a.use_phi0 ? cp0 : us;
          rhs = T(0);
          if (fabs(cur) >= a.zero_tol) {
            const T n2 = FASTW ? godunov_star_sq_undivided<T>(cur, xp, yp, zp, xm, ym, zm, ivx, ivy, ivz)
                               : godunov_star_sq_max3<T>(cur, xp, yp, zp, xm, ym, zm);
            if constexpr (FASTW) {
              // FP32: the smoothed sign through the hardware reciprocal square root
              // and |grad phi|* through the hardware square root (about 1 ulp each)
              const T sgn = a.use_phi0 ? cur * __builtin_amdgcn_rsqf(fma(cur, cur, a.dx_sq))
                                       : cur * __builtin_amdgcn_rsqf(fma(n2, a.dx_sq, cur * cur));
              rhs = sgn * (T(1) - __builtin_amdgcn_sqrtf(n2));
            } else {
              const T sgn = a.use_phi0 ? cur / sqrt(cur * cur + a.dx_sq) : cur / sqrt(cur * cur + n2 * a.dx_sq);
              rhs = sgn * (T(1) - sqrt(n2));
            }
          }
        } else if (KIND == STAGE_ADVECT_LF) {
          // Lax-Friedrichs Hamiltonian on the plus/minus WENO5 derivatives (lf_axis)
          T xp, xm, yp, ym, zp, zm;
          hj_weno5_pm_d<T, FASTW>(dx_, a.ix, tiny_dx2x, xp, xm);
          hj_weno5_pm_d<T, FASTW>(dy_, a.iy, tiny_dx2y, yp, ym);
          hj_weno5_pm_d<T, FASTW>(dz_, a.iz, tiny_dx2z, zp, zm);
          const T hx = lf_axis<T>(cux, a.alpha_mode ? a.alpha[0] : fabs(cux), xp, xm);
          const T hy = lf_axis<T>(cuy, a.alpha_mode ? a.alpha[1] : fabs(cuy), yp, ym);
          const T hz = lf_axis<T>(cuz, a.alpha_mode ? a.alpha[2] : fabs(cuz), zp, zm);
          rhs = -((hx + hy) + hz);
        } else {
          const T gx = upwind_weno5_wave<T, FASTW>(dx_, cux, a.ix, tiny_dx2x, a.zero_tol);
          const T gy = upwind_weno5_wave<T, FASTW>(dy_, cuy, a.iy, tiny_dx2y, a.zero_tol);
          const T gz = upwind_weno5_wave<T, FASTW>(dz_, cuz, a.iz, tiny_dx2z, a.zero_tol);
          // lsm_level_set_evolution3d.f.in:113-116 on a zeroed rhs
          rhs = FASTW ? -fma(cuz, gz, fma(cuy, gy, cux * gx)) : -(cux * gx + cuy * gy + cuz * gz);
        }
        v = FASTW ? rk_blend<T>(uc, us, rhs, a.dt, a.rkb) : rk_combine<T>(a.form, uc, us, rhs, a.dt);
        const bool img_z = IMAGES && (a.periodic & 4) && (k - fb.klo < 3 || fb.khi - k < 3);
        if (img_xy || img_z) store_with_periodic_images<T>(a.out, fb, i, j, k, v, a.periodic, 3);
        else buf_store(v, plane_rsrc(p_out), oo);
      }
      if (f_z && active) {
        // below the bottom face from planes klo (kept) and klo + 1; above the
        // top face from planes khi and khi - 1 (kept)
        if (oc6[4] && k == fb.klo + 1) store_face_ghosts<T>(p_out - s_out, -s_out, oo, 0, oc6[4], v_prev, v);
        if (oc6[5] && k == fb.khi) store_face_ghosts<T>(p_out, s_out, oo, 0, oc6[5], v, v_prev);
      }
      v_prev = v;
      p_out += s_out;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) zv[q] = zv[q + U];
#pragma unroll
    for (int q = 0; q < 5; q++) zd[q] = zd[q + U];
  }
  if (FACES && f_xy && k0 <= k1) {  // uniform per tile
    // x/y face ghosts of every plane of the chunk, after the march: the face
    // lane reads back its own and its inward neighbour's output (stored by
    // this workgroup; the barrier's wait and fence make the stores visible),
    // so the march itself carries no face state in its registers
    __syncthreads();
    if (active) {
      constexpr int32_t e = (int32_t)sizeof(T);
      const int32_t row = (int32_t)a.out.sy * e;
      const char* pl = p_out - (int64_t)(k1 - k0 + 1) * s_out;  // plane k0
      for (int k = k0; k <= k1; k++, pl += s_out) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(pl);
        const T own = buf_load<T>(r, oo);
        if (f_xlo && i == fb.ilo) store_face_ghosts<T>(pl, 0, oo, -e, oc6[0], own, buf_load<T>(r, oo + e));
        if (f_xhi && i == fb.ihi) store_face_ghosts<T>(pl, 0, oo, e, oc6[1], own, buf_load<T>(r, oo - e));
        if (f_ylo && j == fb.jlo) store_face_ghosts<T>(pl, 0, oo, -row, oc6[2], own, buf_load<T>(r, oo + row));
        if (f_yhi && j == fb.jhi) store_face_ghosts<T>(pl, 0, oo, row, oc6[3], own, buf_load<T>(r, oo - row));
      }
    }
  }
}

// Host-side pieces shared by the fused-stage and patch-list entry points
// (defined in lsm_fused.hip).
struct TileShape {
  int tx, ty;
};
struct Tiling {
  int cfg;  // tile shape index
  int kz;   // planes per workgroup
};
Tiling current_tiling();
double current_zero_tol();  // the fused stages' upwind switch threshold (per host thread)
TileShape tile_shape(int cfg);
int tiling_count();
bool valid_form(int form);
bool device_only(const char* fn, const void* p, const char* what);
GhostRule make_ghost_rule(const char* fn, const int* bc, const Box& fb);

// True when no stencil read of tile (bx, by, bz) leaves the fill box, so the
// march needs no ghost values at all (uniform per workgroup).
template <int TX, int TY>
__device__ __forceinline__ bool tile_is_interior(const Box& fb, const TileGrid& tg, int bx, int by, int bz) {
  const int i0 = fb.ilo + bx * TX, j0 = fb.jlo + by * TY, k0 = fb.klo + bz * tg.kz;
  const int k1 = min(k0 + tg.kz - 1, fb.khi);
  return i0 - 3 >= fb.ilo && i0 + TX + 2 <= fb.ihi && j0 - 3 >= fb.jlo && j0 + TY + 2 <= fb.jhi &&
         k0 - 3 >= fb.klo && k1 + 3 <= fb.khi;
}

// Decomposes a logical tile index of one fill box into (bx, by, bz).
__device__ __forceinline__ void tile_coords(const TileGrid& tg, int b, int& bx, int& by, int& bz) {
  bx = b % tg.ntx;
  const int rem = b / tg.ntx;
  by = rem % tg.nty;
  bz = rem / tg.nty;
}

// Tile of workgroup b of a one-tile-per-workgroup launch.  The hardware hands
// workgroup b to XCD b % 8.  When the tile rows split evenly over the 8 XCDs,
// they form 8 slabs of nty/8 rows, and in z-chunk bz XCD x owns slab
// (x + bz) % 8: all XCDs work through the z-chunks together, so the tiles in
// flight cover one z-slab of the grid (HBM pages of few planes at a time; 4 %
// faster than contiguous tile ranges per XCD for the 512^3 advection stage),
// a tile's x and y neighbours stay in its XCD's L2, and the slabs at the y
// faces (whose tiles synthesise or write ghost cells) rotate over the XCDs
// instead of loading two of them.  Otherwise each XCD takes a contiguous
// range of tiles (xcd_remap).  A bijection on [0, count) in both cases.
__device__ __forceinline__ void launch_tile_coords(const TileGrid& tg, int b, int nblk, int& bx, int& by, int& bz) {
  if (tg.nty % kNumXcd == 0) {
    const int x = b % kNumXcd, m = b / kNumXcd;
    const int rows = tg.nty / kNumXcd, per = tg.ntx * rows;
    bz = m / per;
    const int r = m - bz * per;
    by = ((x + bz) % kNumXcd) * rows + r / tg.ntx;
    bx = r % tg.ntx;
  } else {
    tile_coords(tg, xcd_remap(b, nblk), bx, by, bz);
  }
}

}  // namespace lsm
