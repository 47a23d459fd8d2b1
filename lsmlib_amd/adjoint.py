"""Reverse-mode WENO5 / TVD-RK3 advection for torch autograd.

The reference has no adjoint; this module differentiates the fused
advection path of `pylsmlib.Stepper` (the reference sequence
LSM3D_UPWIND_HJ_WENO5 -> LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS ->
LSM3D_TVD_RK3_STAGE1/2/3, LevelSetFunctionIntegrator.cc:1319-1510) on a grid
periodic on every axis:

* `AdvectionRK3` runs nsteps TVD-RK3 steps with the fused stage kernels,
  keeping phi every `checkpoint_every` steps, and runs the reverse sweep with
  the hand-written stage adjoint `lsm3d_amd_advect_stage_adjoint`: each
  segment is recomputed from its checkpoint (storing the stage inputs), then
  walked backwards stage by stage:
      stage 3:  s2bar  = 2/3 (pbar + dt J(s2)^T pbar),   A  = 1/3 pbar
      stage 2:  s1bar  = 1/4 (s2bar + dt J(s1)^T s2bar), A += 3/4 s2bar
      stage 1:  A     += s1bar + dt J(phi)^T s1bar;       pbar <- A
  and V-bar accumulates -a_s dt ubar D_a(u_s) per stage.
* `advect_weno5_rk3(phi0, vel, ...)` is the torch.autograd.Function entry:
  gradients flow to phi0 and to the three velocity components, so a
  velocity built from parameters in torch (for example
  `workloads.FourierVelocity`) chains to those parameters.
* `heaviside_mismatch(phi, target, eps)` is the loss
  sum (H_eps(phi) - H_eps(target))^2 over the fill box with the smoothed
  Heaviside of lsmHeaviside (lsm_calculus_toolbox.f:25-54), computed by a
  device kernel that also returns d loss / d phi.

Fields are (nk, nj, ni) tensors over the grid's ghost box, as in
`pylsmlib`.  The forward reads ghost cells written as periodic images; the
reverse sweep wraps indices inside the fill box, and gradients arriving on
ghost cells of phi_T are folded onto their periodic sources.
"""
import ctypes

import torch

from ._native import toolbox
from .pylsmlib import (BC_PERIODIC, RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, Stepper, _ptr,
                       box_c, real_of)

_ACC_STAGE, _ACC_CUR = 1, 2


def fold_periodic_ghosts(a, g):
    """Adds the values on the ghost layers of `a` to the fill-box cells they
    are periodic images of, and zeroes the ghosts (the transpose of a
    periodic ghost fill).  g: ghost width on every side, or a Grid (per-axis
    widths, x possibly padded).  In place; returns a."""
    if isinstance(g, int):
        widths = {2: (g, g), 1: (g, g), 0: (g, g)}
    else:
        widths = {2: (g.xlo, g.xhi), 1: (g.ghost, g.ghost), 0: (g.ghost, g.ghost)}
    for axis in (2, 1, 0):
        lo, hi = widths[axis]
        n = a.shape[axis] - lo - hi
        if lo > n or hi > n:
            raise ValueError("ghost layers wider than the periodic period")
        lo_ghost = a.narrow(axis, 0, lo)
        hi_ghost = a.narrow(axis, lo + n, hi)
        a.narrow(axis, n, lo).add_(lo_ghost)      # ghost t < 0 is cell t + n
        a.narrow(axis, lo, hi).add_(hi_ghost)     # ghost t >= n is cell t - n
        lo_ghost.zero_()
        hi_ghost.zero_()
    return a


class AdvectionRK3:
    """Forward and reverse WENO5 / TVD-RK3 advection on one periodic grid.

    real: "double" or "float".  The velocity is fixed in time and dt is
    frozen (a CFL-chosen dt is not differentiable).  zero_tol: the upwind
    switch of the forward stages and of their adjoints (|V_a| < zero_tol: no
    advection term along a, so no V_a gradient there); None keeps the
    build's LSMLIB_ZERO_TOL (1e-11 double, 1e-5 float).  A float runner
    compared with a double one wants 1e-11: with 1e-5 the cells where
    1e-11 <= |V_a| < 1e-5 contribute to the double V gradient only."""

    def __init__(self, grid, real="double", device="cuda", checkpoint_every=8, zero_tol=None):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.grid = grid
        self.real = real
        self.stepper = Stepper(grid, real, bc=(BC_PERIODIC,) * 6, device=device, ghosts="images",
                               zero_tol=zero_tol)
        self.lib = self.stepper.lib
        self.dtype = self.stepper.dtype
        self.device = self.stepper.device
        self.every = int(checkpoint_every)
        self._gb = box_c(grid.gb)
        self._fb = box_c(grid.fb)
        n = self.lib.lib.lsm3d_amd_advect_adjoint_work_elems(self._fb)
        self.work = torch.empty(int(n), dtype=self.dtype, device=self.device)

    # -- forward -------------------------------------------------------------
    def _check(self, t, what):
        if t.shape != torch.Size(self.grid.shape) or t.dtype != self.dtype or t.device != self.device:
            raise ValueError(f"{what}: expected {self.dtype} tensor of shape {self.grid.shape} on {self.device}")

    def initial_state(self, phi0):
        """Copy of phi0 with its periodic ghost images written."""
        self._check(phi0, "phi0")
        phi = phi0.detach().clone().contiguous()
        self.stepper.sync(False)
        self.stepper.fill_ghosts(phi)
        return phi

    def run(self, phi0, vel, nsteps, dt, keep_checkpoints=True):
        """phi after nsteps steps, and the checkpoints [phi_0, phi_e, phi_2e, ...]."""
        for c, t in zip("xyz", vel):
            self._check(t, "vel_" + c)
        phi = self.initial_state(phi0)
        s1, s2 = self.stepper.empty(), self.stepper.empty()
        ckpts = []
        for n in range(nsteps):
            if keep_checkpoints and n % self.every == 0:
                ckpts.append(phi.clone())
            self.stepper.advect_rk3_step(phi, s1, s2, vel, dt)
        return phi, ckpts

    # -- reverse ---------------------------------------------------------------
    def stage_adjoint(self, form, ubar, adj_s, adj_c, accumulate, us, vel, vbar, dt):
        dx, dy, dz = self.grid.dx
        gb = self._gb
        self.lib.set_zero_tol(self.stepper.zero_tol)  # the forward stages' switch
        self.lib.fused("lsm3d_amd_advect_stage_adjoint", int(form), _ptr(ubar), gb, _ptr(adj_s), gb,
                       _ptr(adj_c), gb, int(accumulate), _ptr(us), gb, _ptr(vel[0]), _ptr(vel[1]),
                       _ptr(vel[2]), gb,
                       _ptr(vbar[0]) if vbar else None, _ptr(vbar[1]) if vbar else None,
                       _ptr(vbar[2]) if vbar else None, gb, _ptr(self.work), self._fb, dx, dy, dz, dt, 7)

    def reverse(self, ckpts, vel, nsteps, dt, phibar_T, want_vel=True, adjoint_runner=None):
        """Adjoints of phi_0 and (optionally) of V given the adjoint of phi_T.

        adjoint_runner: an AdvectionRK3 of the same grid in another precision
        whose stage adjoints are applied to this runner's recomputed stage
        inputs (cast to its precision), for example FP64 adjoints over FP32
        forward states; the adjoints are then in its precision."""
        st = self.stepper
        st.sync(False)
        ar = self if adjoint_runner is None else adjoint_runner
        if ar.grid is not self.grid:
            raise ValueError("adjoint_runner must be built on the same grid")
        zeros = lambda: torch.zeros(self.grid.shape, dtype=ar.dtype, device=self.device)
        cast = (lambda t: t) if ar.dtype == self.dtype else (lambda t: t.to(ar.dtype))
        pbar = fold_periodic_ghosts(phibar_T.detach().to(ar.dtype).clone().contiguous(), self.grid)
        acc, s1b, s2b = zeros(), zeros(), zeros()
        vbar = [zeros(), zeros(), zeros()] if want_vel else None
        avel = tuple(cast(v).contiguous() for v in vel)
        nseg = (nsteps + self.every - 1) // self.every
        if len(ckpts) != nseg:
            raise ValueError("checkpoints do not match nsteps")
        seg_len = min(self.every, nsteps)
        phis = [st.empty() for _ in range(seg_len)]
        s1s = [st.empty() for _ in range(seg_len)]
        s2s = [st.empty() for _ in range(seg_len)]
        for seg in reversed(range(nseg)):
            n0 = seg * self.every
            m = min(self.every, nsteps - n0)
            phis[0].copy_(ckpts[seg])
            for q in range(m):  # recompute the stage inputs of the segment
                st.advect_stage(RK1, s1s[q], phis[q], None, vel, dt)
                st.advect_stage(TVDRK3_STAGE2, s2s[q], s1s[q], phis[q], vel, dt)
                if q + 1 < m:
                    st.advect_stage(TVDRK3_STAGE3, phis[q + 1], s2s[q], phis[q], vel, dt)
            for q in reversed(range(m)):
                ar.stage_adjoint(TVDRK3_STAGE3, pbar, s2b, acc, 0, cast(s2s[q]), avel, vbar, dt)
                ar.stage_adjoint(TVDRK3_STAGE2, s2b, s1b, acc, _ACC_CUR, cast(s1s[q]), avel, vbar, dt)
                ar.stage_adjoint(RK1, s1b, acc, None, _ACC_STAGE, cast(phis[q]), avel, vbar, dt)
                pbar, acc = acc, pbar
        return pbar, vbar


class _AdvectWeno5RK3(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, vx, vy, vz, runner, nsteps, dt):
        vel = (vx.detach(), vy.detach(), vz.detach())
        phi, ckpts = runner.run(phi0.detach(), vel, nsteps, dt, keep_checkpoints=True)
        ctx.runner, ctx.nsteps, ctx.dt, ctx.ckpts = runner, nsteps, dt, ckpts
        ctx.save_for_backward(vx, vy, vz)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        vx, vy, vz = ctx.saved_tensors
        want_vel = any(ctx.needs_input_grad[1:4])
        pbar, vbar = ctx.runner.reverse(ctx.ckpts, (vx.detach(), vy.detach(), vz.detach()), ctx.nsteps,
                                        ctx.dt, gphi, want_vel=want_vel)
        if vbar is None:
            vbar = [None, None, None]
        return pbar, vbar[0], vbar[1], vbar[2], None, None, None


def advect_weno5_rk3(phi0, vel, runner, nsteps, dt):
    """phi after nsteps WENO5 / TVD-RK3 advection steps, differentiable with
    respect to phi0 and vel = (vx, vy, vz); runner is an AdvectionRK3 of the
    grid and precision of the tensors."""
    if real_of(phi0) != runner.real:
        raise TypeError("phi0 dtype does not match the runner's precision")
    return _AdvectWeno5RK3.apply(phi0, vel[0], vel[1], vel[2], runner, int(nsteps), float(dt))


class _HeavisideMismatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, target, grid, eps):
        lib = toolbox(real_of(phi))
        if target.dtype != phi.dtype or target.shape != phi.shape:
            raise ValueError("target must match phi")
        loss = torch.empty(1, dtype=torch.float64, device=phi.device)
        grad = torch.zeros_like(phi)
        gb, fb = box_c(grid.gb), box_c(grid.fb)
        lib.set_stream(torch.cuda.current_stream(phi.device).cuda_stream)
        lib.set_sync(False)
        lib.fused("lsm3d_amd_heaviside_mismatch", _ptr(phi.detach().contiguous()), gb,
                  _ptr(target.detach().contiguous()), gb, fb, float(eps),
                  ctypes.c_void_p(loss.data_ptr()), _ptr(grad), gb)
        ctx.save_for_backward(grad)
        return loss[0]

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return grad * g.to(grad.dtype), None, None, None


def heaviside_mismatch(phi, target, grid, eps):
    """sum over the fill box of (H_eps(phi) - H_eps(target))^2 (FP64 scalar),
    differentiable with respect to phi."""
    return _HeavisideMismatch.apply(phi, target, grid, eps)
