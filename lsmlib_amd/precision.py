"""Where the FP32 path's error against FP64 comes from, on a given workload.

`precision_split` advances four copies of the same periodic problem in
lockstep with the fused WENO5 / TVD-RK3 stages (and, optionally, periodic
reinitialization with the sign from the current phi):

  ref64   FP64 library, FP64 inputs (the reference run)
  in64    FP64 library, inputs rounded once to FP32 (phi0, velocity, dt)
  st64    as in64, and every stage output rounded to FP32: FP32 storage
          with FP64 arithmetic
  ph64    as in64, with only phi (each step's and each reinitialization
          iteration's final stage output) rounded to FP32
  sg64    as in64, with only the intermediate stage fields s1 and s2
          rounded to FP32
  fp32    FP32 library (FP32 storage and arithmetic), the same FP32 inputs

and reports the near-interface (|phi_ref| < 3 dx) maximum differences in
units of dx:

  fp32_vs_fp64      |fp32 - ref64|, the FP32 path's error
  inputs_only       |in64 - ref64|: a perturbation of half an FP32 ulp of the
                    inputs, grown by the problem itself
  storage_only      |st64 - in64|: what rounding the stored fields to FP32
                    costs with exact FP64 arithmetic
  storage_phi_only  |ph64 - in64|: rounding phi alone (the RK base state)
  storage_stages_only  |sg64 - in64|: rounding s1 and s2 alone
  fp32_same_inputs  |fp32 - in64|: storage plus FP32 arithmetic
  one_step          the FP32 step's local error, one step from the final
                    in64 state rounded to FP32

If storage_only is already above a tolerance, no FP32-storage scheme meets
it, whatever the arithmetic (FP64 weight denominators included).
"""
import time

import torch

from .pylsmlib import RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, Stepper, stable_advection_dt
from .workloads import line_elems, sphere_union_taylor_green

_FORMS = ((RK1, "s1", "phi", None), (TVDRK3_STAGE2, "s2", "s1", "phi"), (TVDRK3_STAGE3, "phi", "s2", "phi"))


class _Run:
    def __init__(self, real, n, round_inputs, rounded, zero_tol, device, velocity):
        dtype = torch.float32 if real == "float" else torch.float64
        grid, phi, vel, bc = sphere_union_taylor_green(n, dtype=torch.float64, device=device,
                                                       align=line_elems(dtype), velocity=velocity)
        dt = stable_advection_dt(vel, grid, 0.5, "double")
        if round_inputs:
            phi = phi.float().double()
            vel = tuple(v.float().double() for v in vel)
            dt = float(torch.tensor(dt, dtype=torch.float32))
        self.grid, self.dt, self.rounded = grid, dt, frozenset(rounded)  # fields stored as FP32
        self.f = {"phi": phi.to(dtype).clone()}
        del phi
        self.vel = tuple(v.to(dtype).contiguous() for v in vel)
        self.st = Stepper(grid, real, bc, zero_tol=zero_tol)
        self.st.fill_ghosts(self.f["phi"])
        self.f["s1"], self.f["s2"] = self.st.empty(), self.st.empty()

    def _stored(self, name):
        if name in self.rounded:
            t = self.f[name]
            t.copy_(t.float().double())

    def step(self):
        for form, o, s, c in _FORMS:
            self.st.advect_stage(form, self.f[o], self.f[s], self.f[c] if c else None, self.vel, self.dt)
            self._stored(o)

    def reinit(self):
        h = 0.5 * self.grid.dx[0]
        for form, o, s, c in _FORMS:
            self.st.reinit_stage(form, self.f[o], self.f[s], self.f[c] if c else None, None, False, h)
            self._stored(o)

    def interior(self):
        return self.grid.interior(self.f["phi"]).double()


def _near_max(a, b, ref, h):
    near = ref.abs() < 3 * h
    e = (a - b).abs()
    e = torch.where(near, e, torch.zeros_like(e))
    return {"max_over_dx": float(e.max()) / h, "cells_above_1e-5_dx": int((e > 1e-5 * h).sum())}


def precision_split(n=512, steps=25, reinit_every=0, reinit_iters=5, zero_tol=1e-11, device="cuda",
                    velocity="taylor_green"):
    """The split above for `steps` steps at n^3 (sphere union, 64 spheres,
    seed 5, Taylor-Green velocity by default), with `reinit_iters`
    reinitialization iterations every `reinit_every` steps (0: none)."""
    t0 = time.perf_counter()
    runs = {"ref64": _Run("double", n, False, (), None, device, velocity),
            "in64": _Run("double", n, True, (), None, device, velocity),
            "st64": _Run("double", n, True, ("phi", "s1", "s2"), None, device, velocity),
            "ph64": _Run("double", n, True, ("phi",), None, device, velocity),
            "sg64": _Run("double", n, True, ("s1", "s2"), None, device, velocity),
            "fp32": _Run("float", n, False, (), zero_tol, device, velocity)}
    for s in range(1, steps + 1):
        for r in runs.values():
            r.step()
        if reinit_every and s % reinit_every == 0:
            for _ in range(reinit_iters):
                for r in runs.values():
                    r.reinit()
    torch.cuda.synchronize()
    h = runs["ref64"].grid.dx[0]
    ref = runs["ref64"].interior()
    vals = {k: r.interior() for k, r in runs.items() if k != "ref64"}
    out = {"fp32_vs_fp64": _near_max(vals["fp32"], ref, ref, h),
           "inputs_only": _near_max(vals["in64"], ref, ref, h),
           "storage_only": _near_max(vals["st64"], vals["in64"], ref, h),
           "storage_phi_only": _near_max(vals["ph64"], vals["in64"], ref, h),
           "storage_stages_only": _near_max(vals["sg64"], vals["in64"], ref, h),
           "fp32_same_inputs": _near_max(vals["fp32"], vals["in64"], ref, h)}
    # the local error of one FP32 step from the final in64 state (rounded to FP32)
    start = vals["in64"].float().double()
    base = runs["in64"]
    for name in ("in64", "fp32"):
        r = runs[name]
        r.grid.interior(r.f["phi"]).copy_(start)
        r.st.fill_ghosts(r.f["phi"])
        r.step()
    torch.cuda.synchronize()
    out["one_step"] = _near_max(runs["fp32"].interior(), base.interior(), base.interior(), h)
    out["workload"] = (f"{n}^3 periodic sphere union (64 spheres) + {velocity.replace('_', '-')}, {steps} WENO5/TVD-RK3 "
                       f"steps" + (f", {reinit_iters} reinitialization iterations every {reinit_every}" if reinit_every
                                   else ""))
    out["seconds"] = time.perf_counter() - t0
    del runs, ref, vals, start
    torch.cuda.empty_cache()
    return out


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def gradient_precision_split(n=128, nsteps=40, every=8, device="cuda"):
    """Where the FP32 gradient's error against FP64 comes from, on the
    gradient workload of `workloads.adjoint_case` (n^3 periodic [-1,1]^3,
    phi0 = |x| - 0.3, 64 x 3 Fourier velocity coefficients c, dt frozen at
    the CFL-0.5 dt, nsteps WENO5 / TVD-RK3 steps, loss
    sum (H_eps(phi_T) - H_eps(phi_target))^2 with eps = 1.5 dx and the target
    advected with c + delta, delta ~ N(0, 0.02^2), seed 9).  Normwise relative
    errors of dL/dc and dL/dphi0 against the FP64 gradient:

      fp32                      FP32 forward and adjoint, upwind switch 1e-11
                                (the double build's)
      fp32_switch_1e-5          the same with the float build's switch 1e-5:
                                cells with 1e-11 <= |V_a| < 1e-5 then carry no
                                V_a gradient, which only the FP64 run has
      fp64_adjoint_fp32_states  FP64 stage adjoints over the FP32 forward
                                states: what FP32 storage of the forward costs
      fp64_fp32_inputs          FP64 throughout from phi0, V and dt rounded to
                                FP32: the problem's own sensitivity
    """
    from .adjoint import AdvectionRK3, heaviside_mismatch
    from .workloads import adjoint_case
    import numpy as np
    dev = torch.device(device)
    grid, phi0, fv, _ = adjoint_case(n, torch.float64, dev)
    c0 = fv.c0.to(dev)
    v0 = fv.velocity(c0)
    dt = stable_advection_dt(tuple(v.contiguous() for v in v0), grid, 0.5, "double")
    del v0
    rng = np.random.default_rng(9)
    delta = torch.tensor(rng.normal(0.0, 0.02, size=tuple(c0.shape)), dtype=torch.float64, device=dev)
    eps = 1.5 * grid.dx[0]
    r64 = AdvectionRK3(grid, "double", dev, every)

    def grads(runner, round_inputs=False, adjoint_runner=None):
        dtype = runner.dtype
        rnd = (lambda t: t.float().double()) if round_inputs else (lambda t: t)
        step = float(np.float32(dt)) if round_inputs else dt
        c = c0.clone().requires_grad_(True)
        v64 = fv.velocity(c)
        vel = tuple(rnd(v.detach()).to(dtype).contiguous() for v in v64)
        with torch.no_grad():
            vt = tuple(rnd(v).to(dtype).contiguous() for v in fv.velocity(c0 + delta))
            p0 = rnd(phi0).to(dtype).contiguous()
            target, _ = runner.run(p0, vt, nsteps, step, keep_checkpoints=False)
            del vt
            phi_t, ckpts = runner.run(p0, vel, nsteps, step)
        pt = phi_t.clone().requires_grad_(True)
        heaviside_mismatch(pt, target, grid, eps).backward()
        with torch.no_grad():
            pbar, vbar = runner.reverse(ckpts, vel, nsteps, step, pt.grad, adjoint_runner=adjoint_runner)
        torch.autograd.backward(list(v64), [b.double() for b in vbar])  # the chain to c in FP64
        torch.cuda.synchronize()
        return c.grad.double(), pbar.double()

    t0 = time.perf_counter()
    ref = grads(r64)
    r32 = AdvectionRK3(grid, "float", dev, every, zero_tol=1e-11)
    runs = {"fp32": lambda: grads(r32),
            "fp32_switch_1e-5": lambda: grads(AdvectionRK3(grid, "float", dev, every)),
            "fp64_adjoint_fp32_states": lambda: grads(r32, adjoint_runner=r64),
            "fp64_fp32_inputs": lambda: grads(r64, round_inputs=True)}
    out = {}
    for name, fn in runs.items():
        gc, gp = fn()
        out[name] = {"c": _rel(gc, ref[0]), "phi0": _rel(gp, ref[1])}
    out["workload"] = (f"{n}^3 periodic [-1,1]^3, phi0 = |x| - 0.3, 64 x 3 Fourier modes, {nsteps} WENO5/TVD-RK3 "
                       f"steps, checkpoint every {every}, Heaviside mismatch loss eps 1.5 dx")
    out["seconds"] = time.perf_counter() - t0
    torch.cuda.empty_cache()
    return out
