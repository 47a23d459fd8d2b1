"""Whole time steps composed from the oracle's toolbox routines, in the order
LSMLIB's integrators call them.  TEST INFRASTRUCTURE ONLY (tests/ and
bench.py's cpu_baseline leg).

  advection stage: LSM3D_UPWIND_HJ_WENO5 -> LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS ->
                   LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS -> LSM3D_TVD_RK3_STAGE*
                   (LevelSetFunctionIntegrator.cc:1319-1510, 1518-1774)
  reinit stage:    LSM3D_HJ_WENO5 -> LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS ->
                   LSM3D_TVD_RK3_STAGE* (ReinitializationAlgorithm.cc:698-1020)
Ghost fill between stages: the LSM3D_*_EXTRAPOLATION routines face by face
(bc codes 1 copy, 2 linear, 3 signed linear) or a periodic copy (code 5; the
reference's periodic fill is SAMRAI's, so it is written out here).
Arrays are numpy (nk, nj, ni) over grid.gb.
"""
import numpy as np

BC_FN = {1: "lsm3dhomogeneousneumannweno5_", 2: "lsm3dlinearextrapolation_",
         3: "lsm3dsignedlinearextrapolation_"}


def periodic_fill_axis(a, g, axis):
    """Ghost layers of width g along one axis (numpy axis 2 = x) from the
    periodic images."""
    n = a.shape[axis] - 2 * g
    sl = [slice(None)] * 3
    src = [slice(None)] * 3
    sl[axis], src[axis] = slice(0, g), slice(n, n + g)
    a[tuple(sl)] = a[tuple(src)]
    sl[axis], src[axis] = slice(n + g, n + 2 * g), slice(g, 2 * g)
    a[tuple(sl)] = a[tuple(src)]


def fill_ghosts(ora, a, grid, bc):
    """Faces 0..5 in order (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi)."""
    g = grid.ghost
    for face in range(6):
        code = bc[face]
        if code == 0:
            continue
        if code == 5:
            if face % 2 == 0:
                periodic_fill_axis(a, g, 2 - face // 2)
            continue
        getattr(ora, BC_FN[code])(a, grid.gb, grid.fb, face)


class OracleStepper:
    def __init__(self, ora, grid, bc):
        self.ora, self.grid, self.bc = ora, grid, tuple(bc)
        z = np.zeros(grid.shape, dtype=ora.dtype)
        self.gx, self.gy, self.gz = z.copy(), z.copy(), z.copy()
        self.xm, self.ym, self.zm = z.copy(), z.copy(), z.copy()
        self.rhs, self.D1 = z.copy(), z.copy()

    def advect_rhs(self, ps, vel):
        o, g = self.ora, self.grid
        dx, dy, dz = g.dx
        o.lsm3dupwindhjweno5_(self.gx, self.gy, self.gz, g.gb, ps, g.gb, vel[0], vel[1], vel[2], g.gb,
                              self.D1, g.gb, g.fb, dx, dy, dz)
        o.lsm3dzerooutlevelseteqnrhs_(self.rhs, g.gb)
        o.lsm3daddadvectiontermtolserhs_(self.rhs, g.gb, self.gx, self.gy, self.gz, g.gb,
                                         vel[0], vel[1], vel[2], g.gb, g.fb)
        return self.rhs

    def reinit_rhs(self, ps, phi0, use_phi0):
        o, g = self.ora, self.grid
        dx, dy, dz = g.dx
        o.lsm3dhjweno5_(self.gx, self.gy, self.gz, g.gb, self.xm, self.ym, self.zm, g.gb, ps, g.gb,
                        self.D1, g.gb, g.fb, dx, dy, dz)
        p0 = phi0 if phi0 is not None else ps
        o.lsm3dcomputereinitializationeqnrhs_(self.rhs, g.gb, ps, g.gb, p0, g.gb,
                                              self.gx, self.gy, self.gz, g.gb,
                                              self.xm, self.ym, self.zm, g.gb, g.fb, dx, dy, dz,
                                              1 if use_phi0 else 0)
        return self.rhs

    def _rk3(self, phi, rhs_of, dt):
        o, g = self.ora, self.grid
        s1, s2, new = phi.copy(), phi.copy(), phi.copy()
        o.lsm3dtvdrk3stage1_(s1, g.gb, phi, g.gb, rhs_of(phi), g.gb, g.fb, dt)
        fill_ghosts(o, s1, g, self.bc)
        o.lsm3dtvdrk3stage2_(s2, g.gb, s1, g.gb, phi, g.gb, rhs_of(s1), g.gb, g.fb, dt)
        fill_ghosts(o, s2, g, self.bc)
        o.lsm3dtvdrk3stage3_(new, g.gb, s2, g.gb, phi, g.gb, rhs_of(s2), g.gb, g.fb, dt)
        fill_ghosts(o, new, g, self.bc)
        phi[...] = new

    def advect_rk3_step(self, phi, vel, dt):
        self._rk3(phi, lambda p: self.advect_rhs(p, vel), dt)

    def reinit_rk3_iteration(self, phi, dt, phi0=None, use_phi0=False):
        self._rk3(phi, lambda p: self.reinit_rhs(p, phi0, use_phi0), dt)

    def advect_rk2_step(self, phi, vel, dt):
        o, g = self.ora, self.grid
        s1, new = phi.copy(), phi.copy()
        o.lsm3dtvdrk2stage1_(s1, g.gb, phi, g.gb, self.advect_rhs(phi, vel), g.gb, g.fb, dt)
        fill_ghosts(o, s1, g, self.bc)
        o.lsm3dtvdrk2stage2_(new, g.gb, s1, g.gb, phi, g.gb, self.advect_rhs(s1, vel), g.gb, g.fb, dt)
        fill_ghosts(o, new, g, self.bc)
        phi[...] = new


def impose_mask(out, mask, phi):
    """IMPOSE_MASK over the whole array (src/serial/lsm_macros.h:84-91)."""
    out[...] = np.where(mask > phi, mask, phi)


def shrink(box, w):
    return (box[0] + w, box[1] - w, box[2] + w, box[3] - w, box[4] + w, box[5] - w)


class CurvatureFlowOracle:
    """The curvature example's constant-curvature model, composed from the
    toolbox routines of `lib` (the oracle restatement, or the flang build of
    the reference Fortran) in the example's order:
    curvatureModelMedium3dMainLoop and reinitializeMedium3d
    (examples/serial/curvature_example/curvature_model3d.c:41-452, 470-626).

    phi_t + a |grad phi| = b kappa |grad phi|: HJ-ENO2 upwind gradients for the
    normal-speed term, second-order central first and second derivatives for
    the curvature term, TVD-RK2 in time, signed linear extrapolation on all six
    faces after each stage, then the mask ('phi <= mask', i.e. max(mask, phi)).
    Arrays are numpy (nk, nj, ni) over grid.gb.  The host-side scalars (dt,
    dt_corr, the time loop) are formed in double precision."""

    DT_MIN_TO_CORRECT, DT_MIN = 1e-5, 0.001  # curvature_model3d.c:36-37

    def __init__(self, lib, grid, mask, a=0.4, b=0.05, cfl=0.5, do_mask=True):
        if a < 0 or b < 0:
            raise ValueError("the example's model takes a >= 0 and b >= 0")
        self.o, self.grid = lib, grid
        self.a, self.b, self.cfl, self.do_mask = float(a), float(b), float(cfl), bool(do_mask)
        self.mask = mask
        z = np.zeros(grid.shape, dtype=lib.dtype)
        names = ("xp", "yp", "zp", "xm", "ym", "zm", "D1", "D2", "px", "py", "pz",
                 "pxx", "pxy", "pxz", "pyy", "pyz", "pzz", "rhs", "s1", "nxt", "phi0")
        for nme in names:
            setattr(self, nme, z.copy())
        self.fb1, self.fb2 = shrink(grid.gb, 1), shrink(grid.gb, 2)
        dx = grid.dx
        # correction for the parabolic (curvature) term, curvature_model3d.c:95-97
        self.dt_corr = (1.0 / (dx[0] * dx[0]) + 1.0 / (dx[1] * dx[1]) + 1.0 / (dx[2] * dx[2])) * (2.0 * self.b)

    def _bc(self, a):
        g = self.grid
        for face in range(6):
            self.o.lsm3dsignedlinearextrapolation_(a, g.gb, g.fb, face)

    def _eno2(self, phi):
        o, g = self.o, self.grid
        o.lsm3dhjeno2_(self.xp, self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb, phi, g.gb,
                       self.D1, g.gb, self.D2, g.gb, g.fb, *g.dx)

    def _rhs(self, phi, want_dt):
        """lse_rhs of one stage (curvature_model3d.c:146-258 / 298-391); returns
        the hyperbolic dt when want_dt and a > 0."""
        o, g = self.o, self.grid
        self.rhs[...] = 0
        dt = None
        if self.a > 0:
            self._eno2(phi)
            o.lsm3daddconstnormalveltermtolserhs_(self.rhs, g.gb, self.xp, self.yp, self.zp, g.gb,
                                                   self.xm, self.ym, self.zm, g.gb, self.a, g.fb)
            if want_dt:
                dt = o.lsm3dcomputestableconstnormalveldt_(self.a, self.xp, self.yp, self.zp, g.gb,
                                                           self.xm, self.ym, self.zm, g.gb, g.fb, *g.dx, self.cfl)
        if self.b > 0:
            o.lsm3dcentralgradorder2_(self.px, self.py, self.pz, g.gb, phi, g.gb, self.fb1, *g.dx)
            o.lsm3dcentralgradorder2_(self.pxx, self.pxy, self.pxz, g.gb, self.px, g.gb, self.fb2, *g.dx)
            o.lsm3dcentralgradorder2_(self.pxy, self.pyy, self.pyz, g.gb, self.py, g.gb, self.fb2, *g.dx)
            o.lsm3dcentralgradorder2_(self.pxz, self.pyz, self.pzz, g.gb, self.pz, g.gb, self.fb2, *g.dx)
            o.lsm3daddconstcurvtermtolserhs_(self.rhs, g.gb, self.px, self.py, self.pz, g.gb,
                                             self.pxx, self.pxy, self.pxz, self.pyy, self.pyz, self.pzz, g.gb,
                                             self.b, self.fb2)
        return dt

    def _choose_dt(self, dt_hyp, dt_sub, tplot):
        """The time step of curvature_model3d.c:186-278 given the hyperbolic
        dt of LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT (a > 0)."""
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
        """One TVD-RK2 step of the model with ghost fill and mask after each
        stage (curvature_model3d.c:146-414); phi is updated in place (with its
        ghost cells).  dt None: the example's stable step for the interval
        position dt_sub of tplot.  Returns dt."""
        o, g = self.o, self.grid
        dt_hyp = self._rhs(phi, dt is None)
        if dt is None:
            dt = self._choose_dt(dt_hyp, dt_sub, tplot)
        o.lsm3dtvdrk2stage1_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, g.fb, dt)
        self._bc(self.s1)
        if self.do_mask:
            impose_mask(self.s1, self.mask, self.s1)
        self._rhs(self.s1, False)
        o.lsm3dtvdrk2stage2_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, g.fb, dt)
        self._bc(self.nxt)
        if self.do_mask:
            impose_mask(phi, self.mask, self.nxt)
        else:
            phi[...] = self.nxt
        return dt

    def reinitialize(self, phi, tmax_r):
        """reinitializeMedium3d (curvature_model3d.c:484-626): HJ-ENO2, TVD-RK2,
        dt = cfl dx, sign from phi (use_phi0_for_sign = 0)."""
        o, g = self.o, self.grid
        dt_r = self.cfl * g.dx[0]
        t_r = 0.0
        self.phi0[...] = phi
        n = 0
        while t_r < tmax_r:
            for stage in (1, 2):
                src = phi if stage == 1 else self.s1
                self._eno2(src)
                o.lsm3dcomputereinitializationeqnrhs_(self.rhs, g.gb, src, g.gb, self.phi0, g.gb,
                                                      self.xp, self.yp, self.zp, g.gb,
                                                      self.xm, self.ym, self.zm, g.gb, g.fb, *g.dx, 0)
                if stage == 1:
                    o.lsm3dtvdrk2stage1_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, g.fb, dt_r)
                    self._bc(self.s1)
                else:
                    o.lsm3dtvdrk2stage2_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, g.fb, dt_r)
                    self._bc(self.nxt)
                    if self.do_mask:
                        impose_mask(phi, self.mask, self.nxt)
                    else:
                        phi[...] = self.nxt
            t_r = t_r + dt_r
            n += 1
        return n

    def run(self, phi, tmax=20.0, tplot=0.1, emax_stop=0.05, do_reinit=True, max_outer=None, max_steps=None):
        """The example's main loop (curvature_model3d.c:41-452) on phi (in
        place): outer intervals of length tplot, each a sequence of stable
        steps, then reinitialization over 5 dx; stops at tmax, when the max
        change over an interval drops below emax_stop * dx, or when the phase
        volume does.  max_outer / max_steps bound the run for tests.
        Returns the run's statistics."""
        o, g = self.o, self.grid
        dx = g.dx
        eps_stop = emax_stop * dx[0]
        eps = 1.5 * dx[0]
        tmax_r = 5 * dx[0]
        vol_max = o.lsm3dvolumeregionphilessthanzero_(self.mask, g.gb, g.fb, *dx, eps)
        vol_phi = o.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.fb, *dx, eps)
        tplot = tplot if tplot < tmax else tmax
        t, max_abs_err = 0.0, 1000.0
        stats = {"outer": 0, "steps": 0, "reinit_iters": 0, "dt_min": 100.0, "dt_max": 0.0,
                 "vol_max": float(vol_max), "history": []}
        phi_prev = phi.copy()
        while t < tmax and max_abs_err > eps_stop and vol_phi > eps_stop:
            if max_outer is not None and stats["outer"] >= max_outer:
                break
            stats["outer"] += 1
            dt_sub = 0.0
            phi_prev[...] = phi
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
            max_abs_err = o.lsm3dmaxnormdiff_(phi, g.gb, phi_prev, g.gb, g.fb)
            vol_phi = o.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.gb, *dx, eps)
            stats["history"].append((float(t), float(max_abs_err), float(vol_phi)))
            if max_steps is not None and stats["steps"] >= max_steps:
                break
        stats["t"] = float(t)
        stats["max_abs_err"] = float(max_abs_err)
        stats["vol_phi"] = float(vol_phi)
        return stats


class FieldExtensionOracle:
    """FieldExtensionAlgorithm::computeExtensionField on one box, composed from
    the toolbox routines of `lib` (FieldExtensionAlgorithm.cc:195-440,
    1005-1170): the PHI_UPWIND signed unit normal of phi (LSM3D_HJ_WENO5 ->
    LSM3D_PHI_UPWIND_GRAD_F -> LSM3D_COMPUTE_SIGNED_UNIT_NORMAL), then TVD-RK3
    steps of S_t = rhs with rhs = LSM3D_UPWIND_HJ_WENO5(S, velocity = normal)
    -> LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS, dt = cfl * min(dx), ghost fill
    of S (bc) after every stage."""

    def __init__(self, lib, grid, bc, cfl=0.5):
        self.o, self.grid, self.bc, self.cfl = lib, grid, tuple(bc), float(cfl)
        z = np.zeros(grid.shape, dtype=lib.dtype)
        self.pm = [z.copy() for _ in range(6)]
        self.grad = [z.copy() for _ in range(3)]
        self.normal = [z.copy() for _ in range(3)]
        self.sg = [z.copy() for _ in range(3)]
        self.rhs, self.D1 = z.copy(), z.copy()
        self.phi = None
        self.dt = self.cfl * min(grid.dx)

    def normals(self, phi):
        o, g = self.o, self.grid
        xp, yp, zp, xm, ym, zm = self.pm
        o.lsm3dhjweno5_(xp, yp, zp, g.gb, xm, ym, zm, g.gb, phi, g.gb, self.D1, g.gb, g.fb, *g.dx)
        o.lsm3dphiupwindgradf_(*self.grad, g.gb, xp, yp, zp, g.gb, xm, ym, zm, g.gb, phi, g.gb, g.fb)
        o.lsm3dcomputesignedunitnormal_(*self.normal, g.gb, *self.grad, g.gb, phi, g.gb, g.fb, *g.dx)
        self.phi = phi
        return self.normal

    def _rhs(self, S):
        o, g = self.o, self.grid
        o.lsm3dupwindhjweno5_(*self.sg, g.gb, S, g.gb, *self.normal, g.gb, self.D1, g.gb, g.fb, *g.dx)
        o.lsm3dcomputefieldextensioneqnrhs_(self.rhs, g.gb, S, g.gb, self.phi, g.gb, *self.sg, g.gb,
                                            *self.normal, g.gb, g.fb, *g.dx)
        return self.rhs

    def step(self, S):
        o, g = self.o, self.grid
        s1, s2, new = S.copy(), S.copy(), S.copy()
        o.lsm3dtvdrk3stage1_(s1, g.gb, S, g.gb, self._rhs(S), g.gb, g.fb, self.dt)
        fill_ghosts(o, s1, g, self.bc)
        o.lsm3dtvdrk3stage2_(s2, g.gb, s1, g.gb, S, g.gb, self._rhs(s1), g.gb, g.fb, self.dt)
        fill_ghosts(o, s2, g, self.bc)
        o.lsm3dtvdrk3stage3_(new, g.gb, s2, g.gb, S, g.gb, self._rhs(s2), g.gb, g.fb, self.dt)
        fill_ghosts(o, new, g, self.bc)
        S[...] = new

    def extend(self, S, phi, max_iterations=100, stop_tol=None):
        self.normals(phi)
        fill_ghosts(self.o, S, self.grid, self.bc)
        count, delta = 0, None
        while count < max_iterations and (stop_tol is None or delta is None or delta > stop_tol):
            prev = S.copy()
            self.step(S)
            count += 1
            if stop_tol is not None:
                delta = self.o.lsm3dmaxnormdiff_(S, self.grid.gb, prev, self.grid.gb, self.grid.fb)
        return count, delta
