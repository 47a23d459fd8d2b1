"""Curvature-driven interface motion on device-resident fields.

The model of LSMLIB's curvature example,

    phi_t + a |grad phi| = b kappa |grad phi|,   phi <= mask,

with second-order HJ-ENO upwinding of the normal-speed term, second-order
central differences for the mean-curvature term, TVD-RK2 in time, signed
linear extrapolation on all six faces after every stage and periodic
reinitialization.  `CurvatureFlow` sequences the fused entry points of
include/lsm3d_amd.h the way curvatureModelMedium3dMainLoop and
reinitializeMedium3d sequence the toolbox routines
(examples/serial/curvature_example/curvature_model3d.c:41-452, 470-626):

  per step:   normal-speed dt (one reduction) -> stage 1 -> ghost fill + mask
              -> stage 2 (in place) -> ghost fill + mask
  per reinit iteration: ENO2 reinit stage 1 -> ghost fill -> stage 2 (in
              place, masked) -> ghost fill + mask

Every kernel reads and writes device memory only; the host keeps the
reference's scalar time loop (dt, interval bookkeeping) in double precision.
"""
import ctypes

import torch

from ._native import toolbox
from .pylsmlib import RK1, TVDRK2_STAGE2, _ptr, box_c, real_of

SIGNED_LINEAR = 3


class CurvatureFlow:
    """Device-resident driver of the curvature example's model for one grid.

    mask: the masking level set (negative in the admissible region), a
    device tensor over grid.gb, or None (no mask).  a, b: normal speed and
    curvature coefficient (a >= 0, b >= 0 as in the example).  zero_tol as
    for pylsmlib.Stepper (None: the build's LSMLIB_ZERO_TOL).  tiling: "auto"
    applies the autotuned (tile shape, z-chunk) of the curvature stages for
    the precision and grid size (lsmlib_amd.tuning), a (cfg, kz) pair fixes
    it, None leaves the library's current setting."""

    DT_MIN_TO_CORRECT, DT_MIN = 1e-5, 0.001  # curvature_model3d.c:36-37

    def __init__(self, grid, real="double", mask=None, a=0.4, b=0.05, cfl=0.5, device="cuda", zero_tol=None,
                 tiling="auto"):
        if a < 0 or b < 0:
            raise ValueError("the example's model takes a >= 0 and b >= 0")
        self.grid, self.real = grid, real
        self.lib = toolbox(real)
        self.device = torch.device(device)
        self.dtype = torch.float64 if real == "double" else torch.float32
        self.a, self.b, self.cfl = float(a), float(b), float(cfl)
        if mask is not None:
            if real_of(mask) != real or tuple(mask.shape) != grid.shape:
                raise ValueError("mask must be a tensor of the grid's shape and precision")
        self.mask = mask
        self._gb, self._fb = box_c(grid.gb), box_c(grid.fb)
        self._bc = box_c((SIGNED_LINEAR,) * 6)
        self.s1 = self.empty()
        self.raw = self.empty() if mask is not None else None
        self.zero_tol = zero_tol
        if tiling == "auto":
            from .tuning import choose_tiling
            tiling = choose_tiling(real, max(grid.n), "curvature")
        self.tiling = tiling
        dx = grid.dx
        self.dt_corr = (1.0 / (dx[0] * dx[0]) + 1.0 / (dx[1] * dx[1]) + 1.0 / (dx[2] * dx[2])) * (2.0 * self.b)
        self.lib.set_sync(False)
        self.lib.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def empty(self):
        return torch.empty(self.grid.shape, dtype=self.dtype, device=self.device)

    def _configure(self):
        self.lib.set_zero_tol(self.zero_tol)
        if self.tiling is not None:
            self.lib.lib.lsm3d_amd_set_stage_tiling(*(int(v) for v in self.tiling))

    # -- kernels -------------------------------------------------------------
    def curvature_stage(self, form, out, ps, pc, dt, masked=True):
        """One stage: out = RK(pc, ps, -a|grad ps| + b kappa |grad ps|) on the
        fill box, masked (and its unmasked boundary layers kept for the ghost
        fill) when the model has a mask and `masked`."""
        self._configure()
        use_mask = masked and self.mask is not None
        dx, dy, dz = self.grid.dx
        self.lib.fused("lsm3d_amd_curvature_stage", form, _ptr(out), self._gb,
                       _ptr(self.raw) if use_mask else None, self._gb, _ptr(ps), self._gb,
                       _ptr(pc) if pc is not None else None, self._gb,
                       _ptr(self.mask) if use_mask else None, self._gb, self._fb, dx, dy, dz,
                       self.a, self.b, dt)

    def reinit_stage(self, form, out, ps, pc, dt, masked):
        self._configure()
        use_mask = masked and self.mask is not None
        dx, dy, dz = self.grid.dx
        self.lib.fused("lsm3d_amd_reinit_eno2_stage", form, _ptr(out), self._gb,
                       _ptr(self.raw) if use_mask else None, self._gb, _ptr(ps), self._gb,
                       _ptr(pc) if pc is not None else None, self._gb, None, self._gb, 0,
                       _ptr(self.mask) if use_mask else None, self._gb, self._fb, dx, dy, dz, dt)

    def fill_ghosts(self, phi, masked=True):
        """Signed linear extrapolation on all six faces in the serial
        package's order, then the mask on the ghost cells.  After a masked
        stage the fill-box values come from the stage's unmasked copy."""
        use_mask = masked and self.mask is not None
        self.lib.fused("lsm3d_amd_extrapolate_ghosts_masked", _ptr(phi), self._gb,
                       _ptr(self.raw) if use_mask else None, self._gb,
                       _ptr(self.mask) if use_mask else None, self._gb, self._fb, self._bc)

    def normal_dt(self, phi):
        """LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT of HJ-ENO2 derivatives of
        phi (one fused reduction; synchronizes)."""
        self._configure()
        out = self.lib._ct(0)
        dx, dy, dz = self.grid.dx
        self.lib.fused("lsm3d_amd_const_normal_vel_dt_eno2", ctypes.byref(out), _ptr(phi), self._gb, self._fb,
                       dx, dy, dz, self.a, self.cfl)
        return float(out.value)

    # -- the example's loop ------------------------------------------------------
    def _choose_dt(self, dt_hyp, dt_sub, tplot):
        dt = dt_hyp if self.a > 0 else tplot
        if self.b > 0:
            max_h = self.cfl / dt if self.a > 0 else 0.0
            dt = self.cfl / (max_h + self.dt_corr)
        if dt < self.DT_MIN_TO_CORRECT:
            dt = self.DT_MIN
        if dt_sub + dt > tplot:
            dt = tplot - dt_sub
        return float(dt)

    def step(self, phi, dt=None, dt_sub=0.0, tplot=0.1):
        """One TVD-RK2 step on phi in place (ghost cells included; they must
        be valid on entry).  dt None: the example's stable step at interval
        position dt_sub of tplot.  Returns dt."""
        if dt is None:
            dt = self._choose_dt(self.normal_dt(phi) if self.a > 0 else None, dt_sub, tplot)
        self.curvature_stage(RK1, self.s1, phi, None, dt)
        self.fill_ghosts(self.s1)
        self.curvature_stage(TVDRK2_STAGE2, phi, self.s1, phi, dt)
        self.fill_ghosts(phi)
        return dt

    def reinitialize(self, phi, tmax_r):
        """reinitializeMedium3d on phi in place: TVD-RK2 iterations of the
        ENO2 reinitialization equation with dt = cfl dx until tmax_r, the
        mask applied after the second stage.  Returns the iteration count."""
        dt_r = self.cfl * self.grid.dx[0]
        t_r, n = 0.0, 0
        while t_r < tmax_r:
            self.reinit_stage(RK1, self.s1, phi, None, dt_r, masked=False)
            self.fill_ghosts(self.s1, masked=False)
            self.reinit_stage(TVDRK2_STAGE2, phi, self.s1, phi, dt_r, masked=True)
            self.fill_ghosts(phi)
            t_r = t_r + dt_r
            n += 1
        return n

    def run(self, phi, tmax=20.0, tplot=0.1, emax_stop=0.05, do_reinit=True, max_outer=None, max_steps=None):
        """The example's main loop (curvature_model3d.c:41-452) on phi in
        place, with its stopping rules (tmax; max change over an interval
        below emax_stop * dx; phase volume below that).  Returns the run's
        statistics (intervals, steps, reinit iterations, dt range, history
        of (t, max change, volume))."""
        L, g = self.lib, self.grid
        dx = g.dx
        eps_stop = emax_stop * dx[0]
        eps = 1.5 * dx[0]
        tmax_r = 5 * dx[0]
        vol_max = L.lsm3dvolumeregionphilessthanzero_(self.mask, g.gb, g.fb, *dx, eps) if self.mask is not None \
            else float("nan")
        vol_phi = L.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.fb, *dx, eps)
        tplot = tplot if tplot < tmax else tmax
        t, max_abs_err = 0.0, 1000.0
        stats = {"outer": 0, "steps": 0, "reinit_iters": 0, "dt_min": 100.0, "dt_max": 0.0,
                 "vol_max": float(vol_max), "history": []}
        phi_prev = torch.empty_like(phi)
        while t < tmax and max_abs_err > eps_stop and vol_phi > eps_stop:
            if max_outer is not None and stats["outer"] >= max_outer:
                break
            stats["outer"] += 1
            dt_sub = 0.0
            phi_prev.copy_(phi)
            while dt_sub < tplot:
                if max_steps is not None and stats["steps"] >= max_steps:
                    break
                dt = self.step(phi, None, dt_sub, tplot)
                stats["dt_max"] = max(stats["dt_max"], dt)
                stats["dt_min"] = min(stats["dt_min"], dt)
                dt_sub = dt_sub + dt
                stats["steps"] += 1
            t = t + dt_sub
            if do_reinit:
                stats["reinit_iters"] += self.reinitialize(phi, tmax_r)
            max_abs_err = L.lsm3dmaxnormdiff_(phi, g.gb, phi_prev, g.gb, g.fb)
            vol_phi = L.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.gb, *dx, eps)
            stats["history"].append((float(t), float(max_abs_err), float(vol_phi)))
            if max_steps is not None and stats["steps"] >= max_steps:
                break
        stats["t"] = float(t)
        stats["max_abs_err"] = float(max_abs_err)
        stats["vol_phi"] = float(vol_phi)
        return stats
