// Batched advance of a list of uneven patches (SAMRAI-style patch level).
//
// The reference advances a patch level patch by patch, one toolbox call per
// patch and routine (LevelSetMethodToolbox.cc:2446-2592 for the RK stage,
// :106-760 for the derivatives).  Here a plan holds every patch's boxes,
// pool offset and constant velocity on the device; one persistent launch per
// Runge-Kutta stage walks a global list of tiles of all patches, and a
// whole multi-step RK3 advance can be captured once into a hipGraph and
// replayed.
//
// Work distribution: tiles are numbered patch by patch; the 8 XCDs each own
// a contiguous range of the tile list of about equal cost (so a patch's halo
// re-reads stay in one L2), ordered costliest first, and that XCD's
// workgroups take tiles from it through an atomic counter until it is empty
// (a dynamic queue: the tiles' real costs vary more than any static
// estimate).  Each queue entry carries its patch, so a tile needs no search.
//
// Ghost values: the reference fills every patch's boundary ghost cells
// before each stage (LevelSetFunctionIntegrator.cc:1319-1510).  Here the
// stage synthesises, on the tiles that touch a face, the face ghost values
// its stencils read (single-face extrapolation of the fill-box cells, equal
// to what the reference's face-by-face fill leaves there,
// lsm_boundary_conditions3d.f:46-640), so no fill pass and no ghost writes
// run between stages; faces with code 0 read the stored ghost cells.
// lsm3d_amd_patch_plan_fill_ghosts writes the face ghost layers of a pool
// for callers that want them stored.
#include <cstdio>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lsm3d_amd.h"
#include "lsm_common.h"
#include "lsm_march.h"
#include "lsm_runtime.h"
#include "lsm_stencil.h"

namespace lsm {

struct PatchDesc {
  Box fb, gb;
  int64_t offset;  // element offset of the patch's ghost box in every pool
  real_t vel[3];
  TileGrid tg;
};

// Face ghost fill of every patch: workgroup (patch, face) = blockIdx.x
// covers that face's columns with a 64 x 4 thread layout (no index
// division); each column's 3 ghost cells get the single-face extrapolation
// of the two fill-box cells next to the face.  Codes: 0 leave, 1 copy,
// 2 linear, 3 signed linear.
template <typename T>
__global__ void __launch_bounds__(256)
    k_patch_fill_faces(const PatchDesc* __restrict__ desc, T* pool, GhostRule g) {
  const int q = (int)blockIdx.x / 6, face = (int)blockIdx.x % 6;
  const int code = g.code[face];
  if (code == 0) return;
  const PatchDesc d = desc[q];
  const Box& fb = d.fb;
  const Field<T> f = make_field<T>(pool + d.offset, d.gb);
  const int axis = face / 2, hi = face % 2;
  const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;  // the two other axes
  const int flo[3] = {fb.ilo, fb.jlo, fb.klo}, fhi[3] = {fb.ihi, fb.jhi, fb.khi};
  const int bnd = hi ? fhi[axis] : flo[axis], step = hi ? -1 : 1;
  const int tx = (int)threadIdx.x % 64, ty = (int)threadIdx.x / 64;
  for (int v = flo[va] + ty; v <= fhi[va]; v += 4) {
    for (int u = flo[ua] + tx; u <= fhi[ua]; u += 64) {
      int idx[3];
      idx[ua] = u;
      idx[va] = v;
      idx[axis] = bnd;
      const T pb = f.at(idx[0], idx[1], idx[2]);
      idx[axis] = bnd + step;
      const T p1 = f.at(idx[0], idx[1], idx[2]);
#pragma unroll
      for (int dist = 1; dist <= 3; dist++) {
        idx[axis] = bnd - step * dist;
        f.at(idx[0], idx[1], idx[2]) = extrapolate<T>(code, pb, p1, dist);
      }
    }
  }
}

// The queue of one stage launch: XCD x owns entries off[x] .. off[x+1]-1 of
// patch[] / tile[] (patch index and tile number within the patch); ctr[x] is
// its next entry and ctr[8] counts finished workgroups.  The last workgroup
// to finish resets the counters, so every launch (graph replays included)
// starts from zero without a memset.
struct PatchQueue {
  const int* off;    // kNumXcd + 1
  const int* patch;  // ntiles
  const int* tile;   // ntiles
  int* ctr;          // kNumXcd + 1, zero between launches
};

// Workgroup b takes tiles from the queue of XCD b % 8 (the hardware hands
// workgroups to the XCDs round-robin).
template <typename T, int TX, int TY, bool SYNTH>
__global__ void __launch_bounds__(TX* TY)
    k_patch_stage(const PatchDesc* __restrict__ desc, PatchQueue qu, T* out_pool, const T* ps_pool,
                  const T* pc_pool, GhostRule ghost, T ix, T iy, T iz, T dt, T zt, T tdx2x, T tdx2y, T tdx2z,
                  int form) {
  __shared__ StageTile<T, TX, TY> tile;
  __shared__ int s_next;
  const int xcd = (int)blockIdx.x % kNumXcd;
  const int base = qu.off[xcd], count = qu.off[xcd + 1] - base;
  for (;;) {
    __syncthreads();  // every lane has read the previous s_next
    if (threadIdx.x == 0) s_next = atomicAdd(&qu.ctr[xcd], 1);
    __syncthreads();
    const int q = s_next;
    if (q >= count) break;
    const int l = qu.patch[base + q];
    const PatchDesc d = desc[l];
    StageArgs<T> a;
    a.out = make_field<T>(out_pool + d.offset, d.gb);
    a.ps = make_field<const T>(ps_pool + d.offset, d.gb);
    if (pc_pool) a.pc = make_field<const T>(pc_pool + d.offset, d.gb);
    else a.pc.p = nullptr;
    a.vx.p = a.vy.p = a.vz.p = a.phi0.p = nullptr;
    a.cvx = T(d.vel[0]);
    a.cvy = T(d.vel[1]);
    a.cvz = T(d.vel[2]);
    a.fb = d.fb;
    a.tg = d.tg;
    a.ghost = ghost;
    a.ix = ix;
    a.iy = iy;
    a.iz = iz;
    a.dt = dt;
    a.dx_sq = T(0);
    a.zero_tol = zt;
    a.tdx2[0] = tdx2x;
    a.tdx2[1] = tdx2y;
    a.tdx2[2] = tdx2z;
    a.form = form;
    a.rkb = rk_blend_weight<T>(form);
    a.periodic = 0;
    a.use_phi0 = 0;
    int bx, by, bz;
    tile_coords(d.tg, qu.tile[base + q], bx, by, bz);
    if (SYNTH && !tile_is_interior<TX, TY>(d.fb, d.tg, bx, by, bz))
      march_tile<T, TX, TY, STAGE_ADVECT_CONST, (sizeof(T) == 4), true, false>(a, bx, by, bz, tile);
    else
      march_tile<T, TX, TY, STAGE_ADVECT_CONST, (sizeof(T) == 4), false, false>(a, bx, by, bz, tile);
  }
  if (threadIdx.x == 0) {
    __threadfence();  // this workgroup's queue reads precede its count
    if (atomicAdd(&qu.ctr[kNumXcd], 1) == (int)gridDim.x - 1)
      for (int x = 0; x <= kNumXcd; x++) qu.ctr[x] = 0;
  }
}

}  // namespace lsm

// ===========================================================================
// Host side: the plan
// ===========================================================================
struct lsm3d_amd_patch_plan {
  int npatch = 0;
  int cfg = 0, kz = 32, tx = 64, ty = 4;
  int64_t ntiles = 0, cells = 0, pool_elems = 0;
  int grid = 0;  // persistent workgroups, a multiple of 8
  int device = 0;
  std::vector<lsm::PatchDesc> host_desc;
  lsm::PatchDesc* desc = nullptr;
  int* queue = nullptr;  // device: off[9], ctr[9], patch[ntiles], tile[ntiles]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
};

namespace {
using namespace lsm;
typedef real_t R;

void launch_fill(hipStream_t s, const lsm3d_amd_patch_plan* p, R* pool, const GhostRule& g) {
  if (!g.any) return;
  hipLaunchKernelGGL(k_patch_fill_faces<R>, dim3((unsigned)(6 * p->npatch)), dim3(256), 0, s, p->desc, pool, g);
}

template <int TX, int TY>
void launch_patch_t(hipStream_t s, const lsm3d_amd_patch_plan* p, R* out, const R* ps, const R* pc,
                    const GhostRule& g, R ix, R iy, R iz, R dt, int form) {
  const R zt = (R)current_zero_tol();
  // 12 * tiny * dx^2 per axis (the WENO5 epsilon floor in undivided units)
  const R tx2 = R(12) * Tol<R>::tiny / (ix * ix), ty2 = R(12) * Tol<R>::tiny / (iy * iy);
  const R tz2 = R(12) * Tol<R>::tiny / (iz * iz);
  PatchQueue qu;
  qu.off = p->queue;
  qu.ctr = p->queue + (kNumXcd + 1);
  qu.patch = p->queue + 2 * (kNumXcd + 1);
  qu.tile = qu.patch + p->ntiles;
  if (g.any)
    hipLaunchKernelGGL((k_patch_stage<R, TX, TY, true>), dim3((unsigned)p->grid), dim3(TX * TY), 0, s, p->desc, qu,
                       out, ps, pc, g, ix, iy, iz, dt, zt, tx2, ty2, tz2, form);
  else
    hipLaunchKernelGGL((k_patch_stage<R, TX, TY, false>), dim3((unsigned)p->grid), dim3(TX * TY), 0, s, p->desc,
                       qu, out, ps, pc, g, ix, iy, iz, dt, zt, tx2, ty2, tz2, form);
}

// One stage over every patch.
void launch_patch(hipStream_t s, const lsm3d_amd_patch_plan* p, R* out, const R* ps, const R* pc,
                  const GhostRule& g, R dx, R dy, R dz, R dt, int form) {
  const R ix = R(1) / dx, iy = R(1) / dy, iz = R(1) / dz;
  switch (p->cfg) {
    case 0: launch_patch_t<64, 4>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 1: launch_patch_t<64, 8>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 2: launch_patch_t<32, 8>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 3: launch_patch_t<128, 2>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 4: launch_patch_t<128, 4>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 5: launch_patch_t<32, 16>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 6: launch_patch_t<16, 16>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    case 7: launch_patch_t<16, 8>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
    default: launch_patch_t<16, 4>(s, p, out, ps, pc, g, ix, iy, iz, dt, form); break;
  }
}

int occupancy_blocks(int cfg) {
  int n = 0;
  hipError_t e = hipSuccess;
  switch (cfg) {
    case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 64, 4, true>, 256, 0); break;
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 64, 8, true>, 512, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 32, 8, true>, 256, 0); break;
    case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 128, 2, true>, 256, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 128, 4, true>, 512, 0); break;
    case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 32, 16, true>, 512, 0); break;
    case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 16, 16, true>, 256, 0); break;
    case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 16, 8, true>, 128, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_patch_stage<R, 16, 4, true>, 64, 0); break;
  }
  if (e != hipSuccess || n < 1) {
    (void)hipGetLastError();
    n = 2;
  }
  return n;
}

// The queue of the persistent launch.  The tiles, in patch order, are cut
// into eight contiguous runs of equal summed cost, one per XCD, so a patch's
// tiles, which share halo cells, stay in one L2 except where a run ends; each
// run is ordered costliest first (longest processing time first), so the
// dynamic queue ends on the cheapest tiles.  A tile's cost estimate is its
// planes plus the six planes its march loads before the first, weighted up
// when its stencils leave the fill box (ghost values synthesised from the
// rule).  Output: off (kNumXcd + 1), then per entry the patch and the tile
// number within the patch.
void build_queue(const std::vector<PatchDesc>& desc, const std::vector<int64_t>& prefix, int tx, int ty,
                 std::vector<int>& off, std::vector<int>& patch, std::vector<int>& tile) {
  const int npatch = (int)desc.size();
  const int64_t ntiles = prefix[npatch];
  std::vector<double> tcost(ntiles);
  std::vector<int> owner(ntiles);
  double total = 0.0;
  for (int l = 0; l < npatch; l++) {
    const PatchDesc& d = desc[l];
    for (int64_t t = prefix[l]; t < prefix[l + 1]; t++) {
      const int bx = (int)((t - prefix[l]) % d.tg.ntx);
      const int rem = (int)((t - prefix[l]) / d.tg.ntx);
      const int by = rem % d.tg.nty, bz = rem / d.tg.nty;
      const int i0 = d.fb.ilo + bx * tx, j0 = d.fb.jlo + by * ty, k0 = d.fb.klo + bz * d.tg.kz;
      const int k1 = std::min(k0 + d.tg.kz - 1, d.fb.khi);
      const bool interior = i0 - 3 >= d.fb.ilo && i0 + tx + 2 <= d.fb.ihi && j0 - 3 >= d.fb.jlo &&
                            j0 + ty + 2 <= d.fb.jhi && k0 - 3 >= d.fb.klo && k1 + 3 <= d.fb.khi;
      tcost[t] = (double)(k1 - k0 + 1 + 6) * (interior ? 1.0 : 1.25);
      owner[t] = l;
      total += tcost[t];
    }
  }
  std::vector<std::vector<int>> xcd_tiles(kNumXcd);
  double run = 0.0;
  for (int64_t t = 0; t < ntiles; t++) {
    const int x = std::min(kNumXcd - 1, (int)((run + 0.5 * tcost[t]) * kNumXcd / total));
    xcd_tiles[x].push_back((int)t);
    run += tcost[t];
  }
  off.assign(kNumXcd + 1, 0);
  patch.clear();
  tile.clear();
  for (int x = 0; x < kNumXcd; x++) {
    std::vector<int> order(xcd_tiles[x]);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tcost[a] > tcost[b]; });
    for (int t : order) {
      patch.push_back(owner[t]);
      tile.push_back((int)(t - prefix[owner[t]]));
    }
    off[x + 1] = (int)patch.size();
  }
}

bool check_plan(const char* fn, const lsm3d_amd_patch_plan* p) {
  if (!p) {
    raise_error(std::string(fn) + ": null plan");
    return false;
  }
  return true;
}

void free_graph(lsm3d_amd_patch_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  p->graph_steps = 0;
}

}  // namespace

extern "C" {

lsm3d_amd_patch_plan* lsm3d_amd_patch_plan_create(int npatch, const int* fbs, const int* gbs,
                                                  const long long* offsets, long long pool_elems,
                                                  const LSMLIB_REAL* vel, int tiling_cfg, int kz) {
  const char* fn = "lsm3d_amd_patch_plan_create";
  lsm::Call c(fn);
  if (npatch < 1 || !fbs || !gbs || !offsets || !vel) raise_error(std::string(fn) + ": empty or null patch list");
  if (tiling_cfg < 0 || tiling_cfg >= tiling_count() || kz < 1) raise_error(std::string(fn) + ": invalid tiling");
  if (!c.ok()) return nullptr;
  auto* p = new lsm3d_amd_patch_plan();
  p->npatch = npatch;
  p->cfg = tiling_cfg;
  p->kz = kz;
  const TileShape ts = tile_shape(tiling_cfg);
  p->tx = ts.tx;
  p->ty = ts.ty;
  p->pool_elems = pool_elems;
  std::vector<int64_t> prefix(npatch + 1, 0);
  for (int q = 0; q < npatch && c.ok(); q++) {
    PatchDesc d{};
    d.fb = Box{fbs[6 * q], fbs[6 * q + 1], fbs[6 * q + 2], fbs[6 * q + 3], fbs[6 * q + 4], fbs[6 * q + 5]};
    d.gb = Box{gbs[6 * q], gbs[6 * q + 1], gbs[6 * q + 2], gbs[6 * q + 3], gbs[6 * q + 4], gbs[6 * q + 5]};
    d.offset = offsets[q];
    d.vel[0] = vel[3 * q];
    d.vel[1] = vel[3 * q + 1];
    d.vel[2] = vel[3 * q + 2];
    if (d.fb.cells() <= 0) raise_error(std::string(fn) + ": empty patch fill box");
    require_inside(fn, "patch ghost box", d.gb, d.fb, 3, 3, 3);
    if (d.offset < 0 || d.offset + d.gb.cells() > pool_elems)
      raise_error(std::string(fn) + ": patch storage outside the pool");
    if (d.fb.nx() < 2 || d.fb.ny() < 2 || d.fb.nz() < 2)
      raise_error(std::string(fn) + ": patches need 2 cells per axis");
    d.tg = make_tile_grid(d.fb, p->tx, p->ty, kz);
    prefix[q + 1] = prefix[q] + d.tg.count();
    p->cells += d.fb.cells();
    p->host_desc.push_back(d);
  }
  p->ntiles = prefix[npatch];
  if (c.ok()) {
    check_hip(hipGetDevice(&p->device), "hipGetDevice");
    int ncu = 0;
    check_hip(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device), "hipDeviceGetAttribute");
    long long want = (long long)ncu * occupancy_blocks(tiling_cfg);
    if (want > p->ntiles) want = p->ntiles;
    want = ((want + kNumXcd - 1) / kNumXcd) * kNumXcd;
    if (want < kNumXcd) want = kNumXcd;
    p->grid = (int)want;
    if (p->ntiles > INT32_MAX) raise_error(std::string(fn) + ": too many tiles for one plan");
    std::vector<int> off, patch, tile;
    if (c.ok()) build_queue(p->host_desc, prefix, p->tx, p->ty, off, patch, tile);
    if (c.ok()) check_hip(hipMalloc(&p->desc, sizeof(PatchDesc) * npatch), "hipMalloc(patch descriptors)");
    const size_t qn = 2 * (kNumXcd + 1) + 2 * (size_t)p->ntiles;
    if (c.ok()) check_hip(hipMalloc(&p->queue, sizeof(int) * qn), "hipMalloc(tile queue)");
    if (c.ok()) {
      std::vector<int> q(qn, 0);  // off, zeroed counters, patches, tiles
      std::copy(off.begin(), off.end(), q.begin());
      std::copy(patch.begin(), patch.end(), q.begin() + 2 * (kNumXcd + 1));
      std::copy(tile.begin(), tile.end(), q.begin() + 2 * (kNumXcd + 1) + p->ntiles);
      check_hip(hipMemcpy(p->desc, p->host_desc.data(), sizeof(PatchDesc) * npatch, hipMemcpyHostToDevice),
                "hipMemcpy(patch descriptors)");
      check_hip(hipMemcpy(p->queue, q.data(), sizeof(int) * qn, hipMemcpyHostToDevice), "hipMemcpy(tile queue)");
    }
  }
  c.finish();
  if (!c.ok()) {
    if (p->desc) (void)hipFree(p->desc);
    if (p->queue) (void)hipFree(p->queue);
    delete p;
    return nullptr;
  }
  return p;
}

void lsm3d_amd_patch_plan_destroy(lsm3d_amd_patch_plan* p) {
  if (!p) return;
  free_graph(p);
  if (p->desc) (void)hipFree(p->desc);
  if (p->queue) (void)hipFree(p->queue);
  delete p;
}

long long lsm3d_amd_patch_plan_cells(const lsm3d_amd_patch_plan* p) { return p ? p->cells : 0; }
long long lsm3d_amd_patch_plan_tiles(const lsm3d_amd_patch_plan* p) { return p ? p->ntiles : 0; }
int lsm3d_amd_patch_plan_workgroups(const lsm3d_amd_patch_plan* p) { return p ? p->grid : 0; }

void lsm3d_amd_patch_plan_stage(lsm3d_amd_patch_plan* p, int rk_form, LSMLIB_REAL* out_pool,
                                const LSMLIB_REAL* stage_pool, const LSMLIB_REAL* cur_pool, LSMLIB_REAL dx,
                                LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6]) {
  const char* fn = "lsm3d_amd_patch_plan_stage";
  lsm::Call c(fn);
  if (!check_plan(fn, p)) return;
  if (!valid_form(rk_form)) raise_error(std::string(fn) + ": unknown Runge-Kutta form");
  if (!out_pool || !stage_pool) raise_error(std::string(fn) + ": null pool");
  if (out_pool == stage_pool) raise_error(std::string(fn) + ": phi_out may not alias phi_stage");
  device_only(fn, out_pool, "out pool");
  device_only(fn, stage_pool, "stage pool");
  device_only(fn, cur_pool, "cur pool");
  GhostRule g = make_ghost_rule(fn, bc, p->host_desc[0].fb);
  for (int f = 0; f < 6 && bc; f++)
    if (bc[f] == 5) raise_error(std::string(fn) + ": patches take extrapolated or stored ghosts, not periodic");
  if (c.ok()) {
    launch_patch(c.stream(), p, out_pool, stage_pool, cur_pool, g, dx, dy, dz, dt, rk_form);
    c.launched("k_patch_stage");
  }
  c.finish();
}

// Captures nsteps TVD-RK3 steps (3 stage launches each; phi advanced in
// place through s1 and s2) into a hipGraph owned by the plan.
void lsm3d_amd_patch_plan_capture_rk3(lsm3d_amd_patch_plan* p, int nsteps, LSMLIB_REAL* phi_pool,
                                      LSMLIB_REAL* s1_pool, LSMLIB_REAL* s2_pool, LSMLIB_REAL dx, LSMLIB_REAL dy,
                                      LSMLIB_REAL dz, LSMLIB_REAL dt, const int bc[6]) {
  const char* fn = "lsm3d_amd_patch_plan_capture_rk3";
  lsm::Call c(fn);
  if (!check_plan(fn, p)) return;
  if (nsteps < 1) raise_error(std::string(fn) + ": nsteps < 1");
  device_only(fn, phi_pool, "phi pool");
  device_only(fn, s1_pool, "s1 pool");
  device_only(fn, s2_pool, "s2 pool");
  GhostRule g = make_ghost_rule(fn, bc, p->host_desc[0].fb);
  for (int f = 0; f < 6 && bc; f++)
    if (bc[f] == 5) raise_error(std::string(fn) + ": patches take extrapolated or stored ghosts, not periodic");
  if (!c.ok()) {
    c.finish();
    return;
  }
  free_graph(p);
  hipStream_t cs = nullptr;
  check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (c.ok()) {
    for (int s = 0; s < nsteps; s++) {
      launch_patch(cs, p, s1_pool, phi_pool, nullptr, g, dx, dy, dz, dt, RK_EULER);
      launch_patch(cs, p, s2_pool, s1_pool, phi_pool, g, dx, dy, dz, dt, RK3_STAGE2);
      launch_patch(cs, p, phi_pool, s2_pool, phi_pool, g, dx, dy, dz, dt, RK3_STAGE3);
    }
    check_hip(hipStreamEndCapture(cs, &p->graph), "hipStreamEndCapture");
  }
  if (c.ok()) check_hip(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  if (c.ok()) p->graph_steps = nsteps;
  if (cs) (void)hipStreamDestroy(cs);
  c.finish();
}

void lsm3d_amd_patch_plan_fill_ghosts(lsm3d_amd_patch_plan* p, LSMLIB_REAL* pool, const int bc[6]) {
  const char* fn = "lsm3d_amd_patch_plan_fill_ghosts";
  lsm::Call c(fn);
  if (!check_plan(fn, p)) return;
  if (!pool) raise_error(std::string(fn) + ": null pool");
  device_only(fn, pool, "pool");
  GhostRule g = make_ghost_rule(fn, bc, p->host_desc[0].fb);
  for (int f = 0; f < 6 && bc; f++)
    if (bc[f] == 5) raise_error(std::string(fn) + ": patches take extrapolated or stored ghosts, not periodic");
  if (c.ok()) {
    launch_fill(c.stream(), p, pool, g);
    c.launched("k_patch_fill_faces");
  }
  c.finish();
}

void lsm3d_amd_patch_plan_replay(lsm3d_amd_patch_plan* p) {
  const char* fn = "lsm3d_amd_patch_plan_replay";
  lsm::Call c(fn);
  if (!check_plan(fn, p)) return;
  if (!p->exec) raise_error(std::string(fn) + ": no captured graph");
  if (c.ok()) check_hip(hipGraphLaunch(p->exec, c.stream()), "hipGraphLaunch");
  c.finish();
}

}  // extern "C"
