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
