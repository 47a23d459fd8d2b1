"""Field extension off the zero level set on device-resident fields.

S_t + sgn(phi) n . grad S = 0 carries a field S away from the interface
phi = 0 along the signed normal, as LSMLIB's FieldExtensionAlgorithm does
(FieldExtensionAlgorithm.cc:195-440) on one box:

  once:       LSM3D_HJ_WENO5 of phi -> LSM3D_PHI_UPWIND_GRAD_F ->
              LSM3D_COMPUTE_SIGNED_UNIT_NORMAL (LevelSetMethodToolbox.cc:3832-4600,
              unit_normal_type PHI_UPWIND), then the extension velocity
              (lsm3d_amd_field_extension_velocity: the normal where
              LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS is active, 0 elsewhere);
  per step:   one TVD-RK3 step of the fused advection stages with that velocity
              (upwind HJ-WENO5 of S, -V.grad S, RK combination; ghost values of
              S synthesised per face), dt = cfl * min(dx);
  stop:       after max_iterations, or when max |S_new - S_old| over the fill
              box drops to stop_tol (LSM3D_MAX_NORM_DIFF).
"""
import torch

from ._native import toolbox
from .pylsmlib import BC_LINEAR, Stepper, _ptr, box_c


class FieldExtension:
    """Extension of fields off the zero level set of phi on one grid.

    bc: ghost-value codes of the extended field per face (1 copy, 2 linear,
    3 signed linear, 5 periodic); phi needs valid ghost cells (3 layers) when
    normals() runs."""

    def __init__(self, grid, real="double", bc=(BC_LINEAR,) * 6, cfl=0.5, device="cuda", tiling="auto"):
        self.grid, self.real, self.cfl = grid, real, float(cfl)
        self.lib = toolbox(real)
        self.stepper = Stepper(grid, real, bc, device=device, tiling=tiling)
        self.device = self.stepper.device
        e = self.stepper.empty
        self.pm = [e() for _ in range(6)]        # plus / minus derivatives of phi
        self.grad = [torch.zeros_like(self.pm[0]) for _ in range(3)]
        self.normal = [e() for _ in range(3)]
        self.vel = [e() for _ in range(3)]
        self.s1, self.s2, self.prev = e(), e(), None
        self.dt = self.cfl * min(grid.dx)

    def normals(self, phi):
        """Signed unit normal of phi (PHI_UPWIND type) and the extension
        velocity; phi must hold valid ghost cells."""
        L, g = self.lib, self.grid
        gb, fb = g.gb, g.fb
        xp, yp, zp, xm, ym, zm = self.pm
        L.lsm3dhjweno5_(xp, yp, zp, gb, xm, ym, zm, gb, phi, gb, self.grad[0], gb, fb, *g.dx)
        L.lsm3dphiupwindgradf_(*self.grad, gb, xp, yp, zp, gb, xm, ym, zm, gb, phi, gb, fb)
        L.lsm3dcomputesignedunitnormal_(*self.normal, gb, *self.grad, gb, phi, gb, fb, *g.dx)
        b = box_c(gb)
        L.fused("lsm3d_amd_field_extension_velocity", *(_ptr(v) for v in self.vel), b,
                *(_ptr(n) for n in self.normal), b, _ptr(phi), b, box_c(fb), *g.dx)
        return self.normal

    def step(self, S):
        """One TVD-RK3 step of the extension equation on S in place."""
        self.stepper.advect_rk3_step(S, self.s1, self.s2, tuple(self.vel), self.dt)

    def extend(self, S, phi, max_iterations=100, stop_tol=None):
        """Extends S (in place, ghost cells valid on entry) off the zero level
        set of phi.  Returns (iterations, last max change or None)."""
        self.normals(phi)
        self.stepper.fill_ghosts(S)
        g = self.grid
        count, delta = 0, None
        if stop_tol is not None and self.prev is None:
            self.prev = self.stepper.empty()
        while count < max_iterations and (stop_tol is None or delta is None or delta > stop_tol):
            if stop_tol is not None:
                self.prev.copy_(S)
            self.step(S)
            count += 1
            if stop_tol is not None:
                delta = self.lib.lsm3dmaxnormdiff_(S, g.gb, self.prev, g.gb, g.fb)
        self.stepper.fill_ghosts(S)
        return count, delta
