"""The curvature example's narrow-band (localized) model on device-resident
fields, sequenced from the drop-in narrow-band routines.

LocalCurvatureFlow runs curvatureModelMedium3dLocalMainLoop and
reinitializeMedium3dLocal (examples/serial/curvature_example/
curvature_model3d_local.c:52-728, 757-934) with every array in device memory:

  per step:  LSM3D_DETERMINE_NARROW_BAND (gamma = 4 dx, beta = 2 dx, levels
             0-3) and the D2 / D1 / ghost-box boundary layers (marks 125 /
             126 / 127, fill-box points <= 124); stage 1 and stage 2 with the
             *LOCAL kernels (ENO2 on levels 0-2, the constant normal speed,
             central first derivatives on levels 0-1 and second derivatives on
             level 0, the curvature term, the cut-off function, TVD-RK2), the
             signed linear ghost fill over the whole box, IMPOSE_MASK_LOCAL
             (entries n_lo[0] .. n_hi[0] - 1, lsm_macros.h:108-117);
  then:      LSM3D_CHECK_OUTER_NARROW_BAND_LAYER and
             LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL decide on reinitialization over
             levels 0+1 (the limits shifted down one level) for gamma + 2 dx.

The stage fields phi_stage1 / phi_next keep, outside the band, what the last
box reinitialization and earlier steps left there, as the reference's do:
the narrow-band stages read them there (the ENO2 stencils of the outer
levels), so this class's initial reinitialization writes them exactly as
reinitializeMedium3d does.

By default a step is one call of the library's native sequence
(lsm3d_amd_local_curvature_step, lsm_local_model.hip), which calls the same
routines in the same order; `run(native=False)` calls them one by one from
here.  Calls are queued on the library stream without host synchronisation; the
host reads back only what the loop branches on (the band's level limits, dt,
the outer-layer sign check, the average gradient, the interval's max
change).  The hyperbolic dt is LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL
with the box routine's |vel_n| and smoothed |grad phi| (the reference's
routine reads two uninitialised variables, lsm_utilities3d_local.f.in:728).
"""
import numpy as np
import torch

import ctypes

from ._native import LocalCurvatureModel, toolbox
from .curvature import CurvatureFlow
from .pylsmlib import _ptr


class LocalCurvatureFlow:
    """Narrow-band driver of the curvature example's model on one grid (the
    arguments as for curvature.CurvatureFlow)."""

    MARK_GB, MARK_D1, MARK_D2, MARK_FB = 127, 126, 125, 124   # curvature_model3d_local.c:43
    LEVEL = 3
    AVE_GRAD_PHI_MIN, AVE_GRAD_PHI_MAX = 0.9, 1.1              # :41-42
    TPLOT = 0.1                                                # curvature_model_top.h:15

    def __init__(self, grid, real="double", mask=None, a=0.4, b=0.05, cfl=0.5, device="cuda", tiling="auto"):
        self.box = CurvatureFlow(grid, real, mask, a, b, cfl, device, tiling=tiling)
        self.grid, self.real, self.mask = grid, real, mask
        self.a, self.b, self.cfl = float(a), float(b), float(cfl)
        self.lib = toolbox(real)
        self.device = self.box.device
        z = lambda: torch.zeros(grid.shape, dtype=self.box.dtype, device=self.device)  # noqa: E731
        # work arrays keep their contents between calls, as the reference's do
        (self.xp, self.yp, self.zp, self.xm, self.ym, self.zm, self.D1, self.D2, self.px, self.py, self.pz,
         self.pxx, self.pxy, self.pxz, self.pyy, self.pyz, self.pzz, self.rhs, self.s1, self.nxt,
         self.phi0) = (z() for _ in range(21))
        n = int(np.prod(grid.shape))
        self.n = n
        self.nb = torch.zeros(grid.shape, dtype=torch.uint8, device=self.device)
        self.ix, self.iy, self.iz = (torch.zeros(n, dtype=torch.int32, device=self.device) for _ in range(3))
        self.outer = torch.zeros(n, dtype=torch.int32, device=self.device)
        self.n_lo = np.zeros(6, dtype=np.int32)
        self.n_hi = np.zeros(6, dtype=np.int32)
        self.outer_limits = (0, -1, 0, -1)
        dx = grid.dx
        self.beta, self.gamma = 2 * dx[0], 4 * dx[0]
        gb = grid.gb
        self.fb1 = (gb[0] + 1, gb[1] - 1, gb[2] + 1, gb[3] - 1, gb[4] + 1, gb[5] - 1)
        self.fb2 = (gb[0] + 2, gb[1] - 2, gb[2] + 2, gb[3] - 2, gb[4] + 2, gb[5] - 2)
        self._model = None

    def model(self):
        """The LsmLocalCurvatureModel of this driver's arrays (for the native
        step, lsm3d_amd_local_curvature_step), with the current limits."""
        m = LocalCurvatureModel()
        for nme in LocalCurvatureModel.ARRAYS:
            attr = {"phi_x_plus": "xp", "phi_y_plus": "yp", "phi_z_plus": "zp", "phi_x_minus": "xm",
                    "phi_y_minus": "ym", "phi_z_minus": "zm", "phi_x": "px", "phi_y": "py", "phi_z": "pz",
                    "phi_xx": "pxx", "phi_xy": "pxy", "phi_xz": "pxz", "phi_yy": "pyy", "phi_yz": "pyz",
                    "phi_zz": "pzz", "lse_rhs": "rhs", "phi_stage1": "s1", "phi_next": "nxt"}.get(nme, nme)
            setattr(m, nme, _ptr(getattr(self, attr)))
        m.mask = _ptr(self.mask) if self.mask is not None else None
        m.narrow_band = _ptr(self.nb)
        m.index_x, m.index_y, m.index_z, m.index_outer = (_ptr(a) for a in (self.ix, self.iy, self.iz, self.outer))
        m.gb[:] = list(self.grid.gb)
        m.fb[:] = list(self.grid.fb)
        m.dx[:] = list(self.grid.dx)
        m.a, m.b, m.cfl = self.a, self.b, self.cfl
        m.n_lo[:] = [int(v) for v in self.n_lo]
        m.n_hi[:] = [int(v) for v in self.n_hi]
        m.outer_limits[:] = [int(v) for v in self.outer_limits]
        return m

    def _pull(self, m):
        self.n_lo[:] = list(m.n_lo)
        self.n_hi[:] = list(m.n_hi)
        self.outer_limits = tuple(m.outer_limits)

    # -- pieces --------------------------------------------------------------------
    def _lists(self, l0, l1):
        return (self.ix, self.iy, self.iz, int(self.n_lo[l0]), int(self.n_hi[l1]))

    def band(self, phi):
        """The narrow band of phi and its boundary-layer marks."""
        L, g = self.lib, self.grid
        self.outer_limits = L.lsm3ddeterminenarrowband_(
            phi, g.gb, self.nb, g.gb, self.ix, self.iy, self.iz, 0, self.n - 1, self.n_lo, self.n_hi, self.outer,
            0, self.n - 1, self.gamma, self.beta, self.LEVEL)
        for box, mark in ((self.fb2, self.MARK_D2), (self.fb1, self.MARK_D1), (g.gb, self.MARK_GB)):
            L.lsm3dmarknarrowbandboundarylayer_(self.nb, g.gb, box, mark)

    def _eno2(self, phi):
        L, g = self.lib, self.grid
        L.lsm3dhjeno2local_(self.xp, self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb, phi, g.gb, self.D1,
                            g.gb, self.D2, g.gb, *g.dx, self.ix, self.iy, self.iz, int(self.n_lo[0]),
                            int(self.n_hi[0]), int(self.n_lo[1]), int(self.n_hi[1]), int(self.n_lo[2]),
                            int(self.n_hi[2]), self.nb, g.gb, self.MARK_FB, self.MARK_D1, self.MARK_D2)

    def _rhs(self, phi, stage):
        L, g = self.lib, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        L.lsm3dzerooutlevelseteqnrhslocal_(self.rhs, g.gb, *self._lists(0, 0 if stage == 1 else 2))
        dt = None
        if self.a > 0:
            self._eno2(phi)
            L.lsm3daddconstnormalveltermtolserhslocal_(self.rhs, g.gb, self.xp, self.yp, self.zp, g.gb, self.xm,
                                                       self.ym, self.zm, g.gb, self.a, *self._lists(0, 0), *nb, fb)
            if stage == 1:
                dt = L.lsm3dcomputestableconstnormalveldtlocal_(
                    self.a, self.xp, self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb, *g.dx, self.cfl,
                    *self._lists(0, 0), *nb, fb)
        if self.b > 0:
            L.lsm3dcentralgradorder2local_(self.px, self.py, self.pz, g.gb, phi, g.gb, *g.dx, *self._lists(0, 1),
                                           *nb, self.MARK_D1)
            for src, outs in ((self.px, (self.pxx, self.pxy, self.pxz)), (self.py, (self.pxy, self.pyy, self.pyz)),
                              (self.pz, (self.pxz, self.pyz, self.pzz))):
                L.lsm3dcentralgradorder2local_(*outs, g.gb, src, g.gb, *g.dx, *self._lists(0, 0), *nb,
                                               self.MARK_D2)
            L.lsm3daddconstcurvtermtolserhslocal_(self.rhs, g.gb, self.px, self.py, self.pz, g.gb, self.pxx,
                                                  self.pxy, self.pxz, self.pyy, self.pyz, self.pzz, g.gb, self.b,
                                                  *self._lists(0, 0), *nb, fb)
        L.lsm3dmultiplycutofflserhslocal_(phi, self.rhs, g.gb, *self._lists(0, 0), self.nb, g.gb, self.MARK_FB,
                                          self.beta, self.gamma)
        return dt

    def _mask(self, dest, src):
        g = self.grid
        if self.mask is None:
            if dest is not src:
                dest.copy_(src)
            return
        self.lib.lsm3dimposemasklocal_(dest, g.gb, src, g.gb, self.mask, g.gb, self.ix, self.iy, self.iz,
                                       int(self.n_lo[0]), int(self.n_hi[0]) - 1)

    def _fill(self, a):
        self.box.fill_ghosts(a, masked=False)

    # -- the loop ------------------------------------------------------------------------
    def step(self, phi, dt_sub=0.0, tplot=0.1):
        """One localized TVD-RK2 step on phi in place; returns dt."""
        L, g = self.lib, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        self._configure()
        self.band(phi)
        dt = self.box._choose_dt(self._rhs(phi, 1), dt_sub, tplot)
        L.lsm3dtvdrk2stage1local_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt, *self._lists(0, 0), *nb, fb)
        self._fill(self.s1)
        if self.mask is not None:
            self._mask(self.s1, self.s1)
        self._rhs(self.s1, 2)
        L.lsm3dtvdrk2stage2local_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt, *self._lists(0, 0),
                                  *nb, fb)
        self._fill(self.nxt)
        if self.mask is not None:
            self._mask(phi, self.nxt)
        else:
            phi.copy_(self.nxt)
        return dt

    def step_native(self, phi, dt_sub=0.0, tplot=0.1):
        """One step through lsm3d_amd_local_curvature_step (the band build and
        the two stages, each fused per level-0 point); returns dt."""
        self._configure()
        m = self._native_model()
        dt = self.lib.fused("lsm3d_amd_local_curvature_step", ctypes.byref(m), _ptr(phi), dt_sub, tplot)
        self._pull(m)
        return dt

    def needs_reinit_native(self, phi, t, dt_sub, last_reinit_time):
        """lsm3d_amd_local_curvature_reinit_trigger after step_native."""
        m = self._native_model()
        code = self.lib.fused("lsm3d_amd_local_curvature_reinit_trigger", ctypes.byref(m), _ptr(phi), t, dt_sub,
                              last_reinit_time)
        return code > 0, ("", "sign", "grad", "time")[code]

    def _native_model(self):
        if self._model is None:
            self._model = self.model()
        return self._model

    def needs_reinit(self, phi, t, dt_sub, last_reinit_time):
        L, g = self.lib, self.grid
        if L.lsm3dcheckouternarrowbandlayer_(phi, g.gb, *self._lists(0, 0), self.outer, 0, self.n - 1,
                                             *self.outer_limits):
            return True, "sign"
        ave = L.lsm3dcomputeavegradphilocal_(phi, g.gb, *g.dx, *self._lists(0, 0), self.nb, g.gb, self.MARK_FB)
        if ave < self.AVE_GRAD_PHI_MIN or ave > self.AVE_GRAD_PHI_MAX:
            return True, "grad"
        return (t + dt_sub - last_reinit_time >= self.TPLOT), "time"

    def reinitialize_local(self, phi, tmax_r):
        """reinitializeMedium3dLocal on the current (shifted) limits."""
        L, g = self.lib, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        dt_r = self.cfl * g.dx[0]
        t_r, n = 0.0, 0
        self.phi0.copy_(phi)
        while t_r < tmax_r:
            for stage in (1, 2):
                src = phi if stage == 1 else self.s1
                self._eno2(src)
                L.lsm3dcomputereinitializationeqnrhslocal_(self.rhs, g.gb, src, g.gb, self.phi0, g.gb, self.xp,
                                                           self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb,
                                                           *g.dx, 0, *self._lists(0, 0), *nb, fb)
                if stage == 1:
                    L.lsm3dtvdrk2stage1local_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt_r, *self._lists(0, 0),
                                              *nb, fb)
                    self._fill(self.s1)
                else:
                    L.lsm3dtvdrk2stage2local_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt_r,
                                              *self._lists(0, 0), *nb, fb)
                    self._fill(self.nxt)
                    if self.mask is not None:
                        self._mask(phi, self.nxt)
                    else:
                        phi.copy_(self.nxt)
            t_r = t_r + dt_r
            n += 1
        return n

    def reinitialize_box(self, phi, tmax_r):
        """reinitializeMedium3d (curvature_model3d.c:470-626) over the whole box,
        with the stage fields in this driver's arrays: the narrow-band loop
        later reads phi_stage1 and phi_next outside the band (points it has
        not updated since), so they must hold what the reference's
        reinitialization leaves there -- the last stage 1 with its ghost fill,
        and the unmasked stage 2 with its ghost fill."""
        from .pylsmlib import RK1, TVDRK2_STAGE2
        m = self.box
        dt_r = self.cfl * self.grid.dx[0]
        t_r, n = 0.0, 0
        self.phi0.copy_(phi)
        while t_r < tmax_r:
            m.reinit_stage(RK1, self.s1, phi, None, dt_r, masked=False)
            m.fill_ghosts(self.s1, masked=False)
            m.reinit_stage(TVDRK2_STAGE2, self.nxt, self.s1, phi, dt_r, masked=False)
            m.fill_ghosts(self.nxt, masked=False)
            if self.mask is not None:
                self.lib.fused("lsm3d_amd_impose_mask", _ptr(phi), _ptr(self.mask), _ptr(self.nxt), phi.numel())
            else:
                phi.copy_(self.nxt)
            t_r = t_r + dt_r
            n += 1
        return n

    def _configure(self):
        self.lib.set_sync(False)
        self.lib.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def run(self, phi, tmax=20.0, tplot=0.1, emax_stop=0.05, max_outer=None, max_steps=None, native=True):
        """curvatureModelMedium3dLocalMainLoop on phi in place (ghost cells
        valid on entry).  native: one library call per step
        (lsm3d_amd_local_curvature_step and its trigger / reinitialization
        companions); False sequences the same routines from Python.  Returns
        the run's statistics."""
        L, g = self.lib, self.grid
        self._configure()
        m = self.model() if native else None
        mref = ctypes.byref(m) if native else None
        pphi = _ptr(phi)
        dx = g.dx
        eps_stop = emax_stop * dx[0]
        eps = 1.5 * dx[0]
        vol_max = (L.lsm3dvolumeregionphilessthanzero_(self.mask, g.gb, g.fb, *dx, eps)
                   if self.mask is not None else float("nan"))
        vol_phi = L.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.fb, *dx, eps)
        tplot = tplot if tplot < tmax else tmax
        stats = {"outer": 0, "steps": 0, "reinits": 0, "reinit_iters": 0, "triggers": {"sign": 0, "grad": 0,
                 "time": 0}, "dt_min": 100.0, "dt_max": 0.0, "vol_max": float(vol_max), "history": [],
                 "band_fraction": []}
        stats["reinit_iters"] += self.reinitialize_box(phi, self.gamma + dx[0])   # :161
        self._configure()
        t, max_abs_err, last_reinit_time = 0.0, 1000.0, 0.0
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
                if native:
                    dt = L.fused("lsm3d_amd_local_curvature_step", mref, pphi, dt_sub, tplot)
                    code = L.fused("lsm3d_amd_local_curvature_reinit_trigger", mref, pphi, t, dt_sub,
                                   last_reinit_time)
                    trigger, why = code > 0, ("", "sign", "grad", "time")[code]
                else:
                    dt = self.step(phi, dt_sub, tplot)
                    trigger, why = self.needs_reinit(phi, t, dt_sub, last_reinit_time)
                stats["dt_max"] = max(stats["dt_max"], dt)
                stats["dt_min"] = min(stats["dt_min"], dt)
                if trigger:
                    stats["triggers"][why] += 1
                    stats["reinits"] += 1
                    last_reinit_time = t
                    if native:
                        stats["reinit_iters"] += L.fused("lsm3d_amd_local_curvature_reinit", mref, pphi,
                                                         self.gamma + 2 * dx[0])
                    else:
                        lo, hi = self.n_lo.copy(), self.n_hi.copy()
                        self.n_hi[0] = self.n_hi[1]
                        self.n_lo[1], self.n_hi[1] = self.n_lo[2], self.n_hi[2]
                        self.n_lo[2], self.n_hi[2] = self.n_lo[3], self.n_hi[3]
                        stats["reinit_iters"] += self.reinitialize_local(phi, self.gamma + 2 * dx[0])
                        self.n_lo[...], self.n_hi[...] = lo, hi
                dt_sub = dt_sub + dt
                stats["steps"] += 1
            t = t + dt_sub
            if native:
                self._pull(m)
            max_abs_err = L.lsm3dmaxnormdifflocal_(phi, g.gb, phi_prev, g.gb, *self._lists(0, 0), self.nb, g.gb,
                                                   self.MARK_FB)
            vol_phi = L.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.gb, *dx, eps)
            stats["band_fraction"].append(float(self.n_hi[2] - self.n_lo[0] + 1) / self.n)
            stats["history"].append((float(t), float(max_abs_err), float(vol_phi)))
            if max_steps is not None and stats["steps"] >= max_steps:
                break
        stats["t"] = float(t)
        stats["max_abs_err"] = float(max_abs_err)
        stats["vol_phi"] = float(vol_phi)
        return stats
