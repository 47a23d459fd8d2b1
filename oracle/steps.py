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


class SlabOracleStepper:
    """OracleStepper's WENO5 / TVD-RK3 advection step with the toolbox
    routines called on k-slabs of the fill box from `threads` host threads:
    the harness-level parallelism a serial toolbox allows.  Per stage every
    slab runs LSM3D_UPWIND_HJ_WENO5 (its own D1 work array),
    LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS on its planes,
    LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS and the TVD-RK3 stage over its fill
    box; after all slabs the ghost fill runs on one thread.  ctypes releases
    the interpreter lock around each library call, so the slabs run
    concurrently.  Each routine writes only its fill box, so the result is
    the serial stepper's bit for bit."""

    def __init__(self, ora, grid, bc, threads):
        from concurrent.futures import ThreadPoolExecutor
        self.ora, self.grid, self.bc = ora, grid, tuple(bc)
        fb, gb = grid.fb, grid.gb
        nk = fb[5] - fb[4] + 1
        self.threads = max(1, min(int(threads), nk))
        cuts = [fb[4] + (nk * q) // self.threads for q in range(self.threads + 1)]
        self.slabs = [(cuts[q], cuts[q + 1] - 1) for q in range(self.threads)]
        z = np.zeros(grid.shape, dtype=ora.dtype)
        self.gx, self.gy, self.gz, self.rhs = z.copy(), z.copy(), z.copy(), z.copy()
        self.D1 = [z.copy() for _ in range(self.threads)]
        self.s1, self.s2, self.new = z.copy(), z.copy(), z.copy()
        self.pool = ThreadPoolExecutor(self.threads)
        self._k0 = gb[4]

    def _slab(self, q, form, out, ps, pc, vel, dt):
        o, g = self.ora, self.grid
        gb, fb = g.gb, g.fb
        k0, k1 = self.slabs[q]
        fbs = (fb[0], fb[1], fb[2], fb[3], k0, k1)
        o.lsm3dupwindhjweno5_(self.gx, self.gy, self.gz, gb, ps, gb, vel[0], vel[1], vel[2], gb, self.D1[q], gb,
                              fbs, *g.dx)
        o.lsm3dzerooutlevelseteqnrhs_(self.rhs[k0 - self._k0:k1 - self._k0 + 1],
                                      (gb[0], gb[1], gb[2], gb[3], k0, k1))
        o.lsm3daddadvectiontermtolserhs_(self.rhs, gb, self.gx, self.gy, self.gz, gb, vel[0], vel[1], vel[2], gb,
                                         fbs)
        if form == 1:
            o.lsm3dtvdrk3stage1_(out, gb, ps, gb, self.rhs, gb, fbs, dt)
        else:
            getattr(o, "lsm3dtvdrk3stage%d_" % form)(out, gb, ps, gb, pc, gb, self.rhs, gb, fbs, dt)

    def _stage(self, form, out, ps, pc, vel, dt):
        futs = [self.pool.submit(self._slab, q, form, out, ps, pc, vel, dt) for q in range(self.threads)]
        for f in futs:
            f.result()
        fill_ghosts(self.ora, out, self.grid, self.bc)

    def advect_rk3_step(self, phi, vel, dt):
        self._stage(1, self.s1, phi, None, vel, dt)
        self._stage(2, self.s2, self.s1, phi, vel, dt)
        self._stage(3, self.new, self.s2, phi, vel, dt)
        phi[...] = self.new

    def close(self):
        self.pool.shutdown()


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


class LocalCurvatureFlowOracle(CurvatureFlowOracle):
    """The curvature example's localized (narrow-band) main loop,
    curvatureModelMedium3dLocalMainLoop and reinitializeMedium3dLocal
    (examples/serial/curvature_example/curvature_model3d_local.c:52-728,
    757-934), composed from the narrow-band routines of `lib`:

      per step:  LSM3D_DETERMINE_NARROW_BAND (width gamma = 4 dx, inner width
                 beta = 2 dx, levels 0-3), the D2 / D1 / ghost-box boundary
                 layers marked 125 / 126 / 127; stage 1 and stage 2 as the box
                 loop but with the *LOCAL kernels on the listed points and the
                 cut-off function; the ghost fill over the whole box; the mask
                 over level 0 (IMPOSE_MASK_LOCAL, lsm_macros.h:108-117: entries
                 n_lo[0] .. n_hi[0] - 1);
      then:      LSM3D_CHECK_OUTER_NARROW_BAND_LAYER; without a sign change,
                 LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL outside [0.9, 1.1], or TPLOT
                 since the last one, triggers reinitialization on levels 0+1
                 (the limits shifted down one level) over gamma + 2 dx;
      interval:  LSM3D_MAX_NORM_DIFF_LOCAL on level 0, the box volume.
    Before the loop phi is reinitialized over the whole box (gamma + dx).
    The hyperbolic dt comes from LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL
    of `dt_lib` (default `lib`): the reference's routine reads two
    uninitialised variables, so a reference build there gives no meaningful
    dt; pass the restatement (which takes the box routine's values)."""

    MARK_GB, MARK_D1, MARK_D2, MARK_FB = 127, 126, 125, 124  # curvature_model3d_local.c:43
    LEVEL = 3
    AVE_GRAD_PHI_MIN, AVE_GRAD_PHI_MAX = 0.9, 1.1            # :41-42
    TPLOT = 0.1                                               # curvature_model_top.h:15

    def __init__(self, lib, grid, mask, a=0.4, b=0.05, cfl=0.5, do_mask=True, dt_lib=None):
        super().__init__(lib, grid, mask, a, b, cfl, do_mask)
        self.dt_lib = dt_lib or lib
        n = int(np.prod(grid.shape))
        self.n = n
        self.nb = np.zeros(grid.shape, dtype=np.uint8)
        self.ix, self.iy, self.iz = (np.zeros(n, dtype=np.int32) for _ in range(3))
        self.n_lo = np.zeros(6, dtype=np.int32)
        self.n_hi = np.zeros(6, dtype=np.int32)
        self.outer = np.zeros(n, dtype=np.int32)
        self.outer_limits = (0, -1, 0, -1)
        self.beta, self.gamma = 2 * grid.dx[0], 4 * grid.dx[0]   # :154

    def _lists(self, l0, l1):
        return (self.ix, self.iy, self.iz, int(self.n_lo[l0]), int(self.n_hi[l1]))

    def band(self, phi):
        o, g = self.o, self.grid
        self.outer_limits = o.lsm3ddeterminenarrowband_(
            phi, g.gb, self.nb, g.gb, self.ix, self.iy, self.iz, 0, self.n - 1, self.n_lo, self.n_hi, self.outer,
            0, self.n - 1, self.gamma, self.beta, self.LEVEL)
        for box, mark in ((self.fb2, self.MARK_D2), (self.fb1, self.MARK_D1), (g.gb, self.MARK_GB)):
            o.lsm3dmarknarrowbandboundarylayer_(self.nb, g.gb, box, mark)

    def _eno2_local(self, phi):
        o, g = self.o, self.grid
        o.lsm3dhjeno2local_(self.xp, self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb, phi, g.gb,
                            self.D1, g.gb, self.D2, g.gb, *g.dx, self.ix, self.iy, self.iz,
                            int(self.n_lo[0]), int(self.n_hi[0]), int(self.n_lo[1]), int(self.n_hi[1]),
                            int(self.n_lo[2]), int(self.n_hi[2]), self.nb, g.gb, self.MARK_FB, self.MARK_D1,
                            self.MARK_D2)

    def _rhs_local(self, phi, stage):
        """lse_rhs of one stage (:231-404 / :431-565); returns the hyperbolic
        dt in stage 1 when a > 0."""
        o, g = self.o, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        o.lsm3dzerooutlevelseteqnrhslocal_(self.rhs, g.gb, *self._lists(0, 0 if stage == 1 else 2))
        dt = None
        if self.a > 0:
            self._eno2_local(phi)
            o.lsm3daddconstnormalveltermtolserhslocal_(self.rhs, g.gb, self.xp, self.yp, self.zp, g.gb, self.xm,
                                                       self.ym, self.zm, g.gb, self.a, *self._lists(0, 0), *nb, fb)
            if stage == 1:
                dt = self.dt_lib.lsm3dcomputestableconstnormalveldtlocal_(
                    self.a, self.xp, self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb, *g.dx, self.cfl,
                    *self._lists(0, 0), *nb, fb)
        if self.b > 0:
            o.lsm3dcentralgradorder2local_(self.px, self.py, self.pz, g.gb, phi, g.gb, *g.dx, *self._lists(0, 1),
                                           *nb, self.MARK_D1)
            for src, outs in ((self.px, (self.pxx, self.pxy, self.pxz)), (self.py, (self.pxy, self.pyy, self.pyz)),
                              (self.pz, (self.pxz, self.pyz, self.pzz))):
                o.lsm3dcentralgradorder2local_(*outs, g.gb, src, g.gb, *g.dx, *self._lists(0, 0), *nb,
                                               self.MARK_D2)
            o.lsm3daddconstcurvtermtolserhslocal_(self.rhs, g.gb, self.px, self.py, self.pz, g.gb, self.pxx,
                                                  self.pxy, self.pxz, self.pyy, self.pyz, self.pzz, g.gb, self.b,
                                                  *self._lists(0, 0), *nb, fb)
        return dt

    def _cutoff(self, phi):
        self.o.lsm3dmultiplycutofflserhslocal_(phi, self.rhs, self.grid.gb, *self._lists(0, 0), self.nb,
                                               self.grid.gb, self.MARK_FB, self.beta, self.gamma)

    def _mask_local(self, dest, src):
        """IMPOSE_MASK_LOCAL (lsm_macros.h:108-117): entries n_lo[0] .. n_hi[0]-1."""
        g = self.grid
        self.o.lsm3dimposemasklocal_(dest, g.gb, src, g.gb, self.mask, g.gb, self.ix, self.iy, self.iz,
                                     int(self.n_lo[0]), int(self.n_hi[0]) - 1)

    def step_local(self, phi, dt_sub, tplot):
        """One localized TVD-RK2 step (:185-570); returns dt."""
        o, g = self.o, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        self.band(phi)
        dt = self._choose_dt(self._rhs_local(phi, 1), dt_sub, tplot)
        self._cutoff(phi)
        o.lsm3dtvdrk2stage1local_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt, *self._lists(0, 0), *nb, fb)
        self._bc(self.s1)
        if self.do_mask:
            self._mask_local(self.s1, self.s1)
        self._rhs_local(self.s1, 2)
        self._cutoff(self.s1)
        o.lsm3dtvdrk2stage2local_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt, *self._lists(0, 0),
                                  *nb, fb)
        self._bc(self.nxt)
        if self.do_mask:
            self._mask_local(phi, self.nxt)
        else:
            phi[...] = self.nxt
        return dt

    def needs_reinit(self, phi, t, dt_sub, last_reinit_time):
        """(:592-652) -> (trigger, reason)"""
        o, g = self.o, self.grid
        change = o.lsm3dcheckouternarrowbandlayer_(phi, g.gb, *self._lists(0, 0), self.outer, 0, self.n - 1,
                                                   *self.outer_limits)
        if change:
            return True, "sign"
        ave = o.lsm3dcomputeavegradphilocal_(phi, g.gb, *g.dx, *self._lists(0, 0), self.nb, g.gb, self.MARK_FB)
        if ave < self.AVE_GRAD_PHI_MIN or ave > self.AVE_GRAD_PHI_MAX:
            return True, "grad"
        return (t + dt_sub - last_reinit_time >= self.TPLOT), "time"

    def reinitialize_local(self, phi, tmax_r):
        """reinitializeMedium3dLocal (:757-934) on the shifted limits."""
        o, g = self.o, self.grid
        nb, fb = (self.nb, g.gb), self.MARK_FB
        dt_r = self.cfl * g.dx[0]
        t_r, n = 0.0, 0
        self.phi0[...] = phi
        while t_r < tmax_r:
            for stage in (1, 2):
                src = phi if stage == 1 else self.s1
                self._eno2_local(src)
                o.lsm3dcomputereinitializationeqnrhslocal_(self.rhs, g.gb, src, g.gb, self.phi0, g.gb, self.xp,
                                                           self.yp, self.zp, g.gb, self.xm, self.ym, self.zm, g.gb,
                                                           *g.dx, 0, *self._lists(0, 0), *nb, fb)
                if stage == 1:
                    o.lsm3dtvdrk2stage1local_(self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt_r, *self._lists(0, 0),
                                              *nb, fb)
                    self._bc(self.s1)
                else:
                    o.lsm3dtvdrk2stage2local_(self.nxt, g.gb, self.s1, g.gb, phi, g.gb, self.rhs, g.gb, dt_r,
                                              *self._lists(0, 0), *nb, fb)
                    self._bc(self.nxt)
                    if self.do_mask:
                        self._mask_local(phi, self.nxt)
                    else:
                        phi[...] = self.nxt
            t_r = t_r + dt_r
            n += 1
        return n

    def run(self, phi, tmax=20.0, tplot=0.1, emax_stop=0.05, max_outer=None, max_steps=None):
        """curvatureModelMedium3dLocalMainLoop on phi (in place).  Returns the
        run's statistics, with the band sizes and reinitialization triggers."""
        o, g = self.o, self.grid
        dx = g.dx
        eps_stop = emax_stop * dx[0]
        eps = 1.5 * dx[0]
        vol_max = o.lsm3dvolumeregionphilessthanzero_(self.mask, g.gb, g.fb, *dx, eps)
        vol_phi = o.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.fb, *dx, eps)
        tplot = tplot if tplot < tmax else tmax
        stats = {"outer": 0, "steps": 0, "reinits": 0, "reinit_iters": 0, "triggers": {"sign": 0, "grad": 0,
                 "time": 0}, "dt_min": 100.0, "dt_max": 0.0, "vol_max": float(vol_max), "history": [],
                 "band_fraction": []}
        stats["reinit_iters"] += self.reinitialize(phi, self.gamma + dx[0])   # :161
        t, max_abs_err, last_reinit_time = 0.0, 1000.0, 0.0
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
                dt = self.step_local(phi, dt_sub, tplot)
                stats["dt_max"] = max(stats["dt_max"], dt)
                stats["dt_min"] = min(stats["dt_min"], dt)
                trigger, why = self.needs_reinit(phi, t, dt_sub, last_reinit_time)
                if trigger:
                    stats["triggers"][why] += 1
                    stats["reinits"] += 1
                    last_reinit_time = t
                    lo, hi = self.n_lo.copy(), self.n_hi.copy()
                    self.n_hi[0] = self.n_hi[1]
                    self.n_lo[1], self.n_hi[1] = self.n_lo[2], self.n_hi[2]
                    self.n_lo[2], self.n_hi[2] = self.n_lo[3], self.n_hi[3]
                    stats["reinit_iters"] += self.reinitialize_local(phi, self.gamma + 2 * dx[0])
                    self.n_lo[...], self.n_hi[...] = lo, hi
                dt_sub = dt_sub + dt
                stats["steps"] += 1
            t = t + dt_sub
            max_abs_err = o.lsm3dmaxnormdifflocal_(phi, g.gb, phi_prev, g.gb, *self._lists(0, 0), self.nb, g.gb,
                                                   self.MARK_FB)
            vol_phi = o.lsm3dvolumeregionphilessthanzero_(phi, g.gb, g.gb, *dx, eps)
            frac = float(self.n_hi[2] - self.n_lo[0] + 1) / self.n
            stats["band_fraction"].append(frac)
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
