"""A level of uneven patches advanced with one launch per Runge-Kutta stage.

`random_patch_level(...)` builds the batched workload: patches with extents
drawn uniformly per axis (seeded), 3 ghost layers, spacing dx, origins
U[0, 1 - n dx], phi = minimum of 1-4 sphere signed distances (centres
U[0,1]^3, radii U[0.05, 0.3]) and a constant velocity U[-1,1]^3 per patch.

`PatchLevel` stores every field of all patches in one pool tensor (each
patch's ghost box contiguous, column-major) and drives the plan API of
include/lsm3d_amd.h: per-stage launches, or a captured hipGraph of several
RK3 steps.  Ghost values come from per-face extrapolation synthesised inside
the stage kernel (default signed linear, as LSM3D_SIGNED_LINEAR_EXTRAPOLATION
fills them between the reference's per-patch stages); `fill_ghosts` writes
them into a pool when they are wanted stored.
"""
import ctypes

import numpy as np
import torch

from ._native import toolbox
from .pylsmlib import BC_SIGNED_LINEAR, RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, box_c

_DT = {"double": torch.float64, "float": torch.float32}


def random_patch_level(npatch=512, lo=16, hi=96, seed=3, dx=1.0 / 128, ghost=3):
    """Patch boxes, velocities and sphere sets (pure description, no arrays)."""
    rng = np.random.default_rng(seed)
    ext = rng.integers(lo, hi + 1, size=(npatch, 3))
    origins = rng.uniform(0.0, 1.0 - ext * dx)
    vel = rng.uniform(-1.0, 1.0, size=(npatch, 3))
    spheres = []
    for _ in range(npatch):
        k = int(rng.integers(1, 5))
        spheres.append((rng.uniform(0, 1, size=(k, 3)), rng.uniform(0.05, 0.3, size=k)))
    return {"ext": ext, "origins": origins, "vel": vel, "spheres": spheres, "dx": dx, "ghost": ghost}


class PatchLevel:
    def __init__(self, desc, real="double", device="cuda", bc=(BC_SIGNED_LINEAR,) * 6, tiling="auto",
                 zero_tol=None, align=0):
        """align (elements, 0 = off): pad each patch's x ghost region and
        round its pool offset so that rows and the first fill-box cell of
        every row start on `align`-element boundaries (32 FP32 / 16 FP64
        = 128-byte cache lines), as pylsmlib.Grid does."""
        self.real = real
        self.zero_tol = zero_tol
        self.align = int(align)
        self.dtype = _DT[real]
        self.device = torch.device(device)
        self.lib = toolbox(real)
        self.desc = desc
        g = desc["ghost"]
        ext = desc["ext"]
        self.npatch = len(ext)
        # index boxes: patch q covers cells 0..n-1 per axis (its own index space)
        self.fbs = np.zeros((self.npatch, 6), dtype=np.int32)
        self.gbs = np.zeros((self.npatch, 6), dtype=np.int32)
        self.xlo = np.zeros(self.npatch, dtype=np.int64)
        self.rowlen = np.zeros(self.npatch, dtype=np.int64)
        al = self.align
        for q, (nx, ny, nz) in enumerate(ext):
            xlo = max(g, al) if al else g
            xhi = g + ((-(nx + xlo + g)) % al if al else 0)
            self.fbs[q] = (0, nx - 1, 0, ny - 1, 0, nz - 1)
            self.gbs[q] = (-xlo, nx - 1 + xhi, -g, ny - 1 + g, -g, nz - 1 + g)
            self.xlo[q], self.rowlen[q] = xlo, nx + xlo + xhi
        sizes = (self.rowlen * (ext[:, 1] + 2 * g) * (ext[:, 2] + 2 * g)).astype(np.int64)
        offsets, o = [], 0
        for n in sizes:
            offsets.append(o)
            o += int(n)
            if al:
                o = -(-o // al) * al
        self.offsets = np.array(offsets, dtype=np.int64)
        self.pool_elems = int(o)
        self.sizes = sizes
        self.cells = int(np.prod(ext, axis=1).sum())
        self.bc = box_c(bc)
        self.dx = desc["dx"]
        ct = ctypes.c_double if real == "double" else ctypes.c_float
        vel = np.ascontiguousarray(desc["vel"], dtype=np.float64 if real == "double" else np.float32)
        self._vel = vel
        if tiling == "auto":  # the autotuned tile shape for patch levels
            from .tuning import choose_tiling
            tiling = choose_tiling(real, 0, "patches")
        cfg, kz = tiling
        self.tiling = (int(cfg), int(kz))
        self.plan = self.lib.fused(
            "lsm3d_amd_patch_plan_create", self.npatch,
            self.fbs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            self.gbs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), self.pool_elems,
            vel.ctypes.data_as(ctypes.POINTER(ct)), int(cfg), int(kz))
        if not self.plan:
            raise RuntimeError("patch plan creation failed")
        self.lib.set_sync(False)
        self.lib.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def __del__(self):
        try:
            if getattr(self, "plan", None):
                self.lib.lib.lsm3d_amd_patch_plan_destroy(self.plan)
        except Exception:
            pass

    # -- fields ------------------------------------------------------------
    def empty_pool(self):
        return torch.empty(self.pool_elems, dtype=self.dtype, device=self.device)

    def patch_view(self, pool, q):
        """Patch q's ghost box of `pool` as (nk, nj, ni)."""
        nx, ny, nz = self.desc["ext"][q]
        g = self.desc["ghost"]
        o, n = int(self.offsets[q]), int(self.sizes[q])
        return pool[o:o + n].view(nz + 2 * g, ny + 2 * g, int(self.rowlen[q]))

    def interior_view(self, pool, q):
        """Patch q's fill box of `pool` as (nk, nj, ni)."""
        nx = int(self.desc["ext"][q][0])
        g, xlo = self.desc["ghost"], int(self.xlo[q])
        v = self.patch_view(pool, q)
        return v[g:v.shape[0] - g, g:v.shape[1] - g, xlo:xlo + nx]

    def initial_phi(self, on_host=False):
        """phi = min over the patch's spheres of the signed distance, at cell
        centres x = origin + (i + 1/2) dx, over the ghost box.  on_host
        builds the pool with numpy and copies it to the device once (one
        copy instead of thousands of small device kernels)."""
        if on_host:
            return self._initial_phi_host()
        pool = torch.zeros(self.pool_elems, dtype=torch.float64, device=self.device)
        h = self.dx
        for q in range(self.npatch):
            o = self.desc["origins"][q]
            gb = self.gbs[q]
            ax = [o[d] + (torch.arange(int(gb[2 * d]), int(gb[2 * d + 1]) + 1, dtype=torch.float64,
                                       device=self.device) + 0.5) * h for d in range(3)]
            Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
            phi = torch.full_like(X, float("inf"))
            for c, r in zip(*self.desc["spheres"][q]):
                phi = torch.minimum(phi, torch.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r)
            self.patch_view(pool, q).copy_(phi)
        return pool.to(self.dtype)

    def _initial_phi_host(self):
        pool = np.zeros(self.pool_elems, dtype=np.float64)
        h = self.dx
        g = self.desc["ghost"]
        for q in range(self.npatch):
            o = self.desc["origins"][q]
            gb = self.gbs[q]
            ax = [o[d] + (np.arange(int(gb[2 * d]), int(gb[2 * d + 1]) + 1, dtype=np.float64) + 0.5) * h
                  for d in range(3)]
            X, Y, Z = ax[0][None, None, :], ax[1][None, :, None], ax[2][:, None, None]
            phi = np.full((ax[2].size, ax[1].size, ax[0].size), np.inf)
            for c, r in zip(*self.desc["spheres"][q]):
                phi = np.minimum(phi, np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r)
            nz, ny = self.desc["ext"][q][2] + 2 * g, self.desc["ext"][q][1] + 2 * g
            assert phi.shape == (nz, ny, int(self.rowlen[q]))
            pool[int(self.offsets[q]):int(self.offsets[q]) + int(self.sizes[q])] = phi.ravel()
        return torch.from_numpy(pool).to(dtype=self.dtype, device=self.device)

    def stable_dt(self, cfl=0.5):
        """min over patches of LSM3D_COMPUTE_STABLE_ADVECTION_DT with the
        patch's constant velocity (lsm_utilities3d.f.in:297-323), the global
        dt the reference takes by a min-reduction over patches
        (LevelSetFunctionIntegrator.cc:567)."""
        np_t = np.float64 if self.real == "double" else np.float32
        inv = np_t(1) / np_t(self.dx)
        v = np.abs(self._vel.astype(np_t))
        m = ((v[:, 0] * inv + v[:, 1] * inv) + v[:, 2] * inv).max()
        tiny = 1e-99 if self.real == "double" else np.float32(1e-35)
        return float(np_t(cfl) / (np_t(max(m, -1.0)) + np_t(tiny)))

    # -- stepping ------------------------------------------------------------
    def stage(self, form, out, ps, pc, dt):
        h = self.dx
        self.lib.set_zero_tol(self.zero_tol)
        self.lib.fused("lsm3d_amd_patch_plan_stage", self.plan, form, ctypes.c_void_p(out.data_ptr()),
                       ctypes.c_void_p(ps.data_ptr()), ctypes.c_void_p(pc.data_ptr()) if pc is not None else None,
                       h, h, h, dt, self.bc)

    def fill_ghosts(self, pool):
        self.lib.fused("lsm3d_amd_patch_plan_fill_ghosts", self.plan, ctypes.c_void_p(pool.data_ptr()), self.bc)

    def rk3_step(self, phi, s1, s2, dt):
        self.stage(RK1, s1, phi, None, dt)
        self.stage(TVDRK3_STAGE2, s2, s1, phi, dt)
        self.stage(TVDRK3_STAGE3, phi, s2, phi, dt)

    def capture_rk3(self, nsteps, phi, s1, s2, dt):
        """Captures `nsteps` TVD-RK3 steps of phi (stage pools s1, s2) into
        the plan's hipGraph; replay() runs them.  The graph holds the pools'
        device addresses, so the level keeps the three tensors alive until the
        next capture."""
        h = self.dx
        self.lib.set_zero_tol(self.zero_tol)
        self.lib.fused("lsm3d_amd_patch_plan_capture_rk3", self.plan, int(nsteps),
                       ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(s1.data_ptr()),
                       ctypes.c_void_p(s2.data_ptr()), h, h, h, dt, self.bc)
        self._graph_pools = (phi, s1, s2)

    def replay(self):
        self.lib.fused("lsm3d_amd_patch_plan_replay", self.plan)

    @property
    def tiles(self):
        return self.lib.lib.lsm3d_amd_patch_plan_tiles(self.plan)

    @property
    def workgroups(self):
        return self.lib.lib.lsm3d_amd_patch_plan_workgroups(self.plan)
