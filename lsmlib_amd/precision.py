"""Where the FP32 path's error against FP64 comes from, on a given workload.

`precision_split` advances four copies of the same periodic problem in
lockstep with the fused WENO5 / TVD-RK3 stages (and, optionally, periodic
reinitialization with the sign from the current phi):

  ref64   FP64 library, FP64 inputs (the reference run)
  in64    FP64 library, inputs rounded once to FP32 (phi0, velocity, dt)
  st64    as in64, and every stage output rounded to FP32: FP32 storage
          with FP64 arithmetic
  fp32    FP32 library (FP32 storage and arithmetic), the same FP32 inputs

and reports the near-interface (|phi_ref| < 3 dx) maximum differences in
units of dx:

  fp32_vs_fp64      |fp32 - ref64|, the FP32 path's error
  inputs_only       |in64 - ref64|: a perturbation of half an FP32 ulp of the
                    inputs, grown by the problem itself
  storage_only      |st64 - in64|: what rounding the stored fields to FP32
                    costs with exact FP64 arithmetic
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
    def __init__(self, real, n, round_inputs, round_storage, zero_tol, device, velocity):
        dtype = torch.float32 if real == "float" else torch.float64
        grid, phi, vel, bc = sphere_union_taylor_green(n, dtype=torch.float64, device=device,
                                                       align=line_elems(dtype), velocity=velocity)
        dt = stable_advection_dt(vel, grid, 0.5, "double")
        if round_inputs:
            phi = phi.float().double()
            vel = tuple(v.float().double() for v in vel)
            dt = float(torch.tensor(dt, dtype=torch.float32))
        self.grid, self.dt, self.round_storage = grid, dt, round_storage
        self.f = {"phi": phi.to(dtype).clone()}
        del phi
        self.vel = tuple(v.to(dtype).contiguous() for v in vel)
        self.st = Stepper(grid, real, bc, zero_tol=zero_tol)
        self.st.fill_ghosts(self.f["phi"])
        self.f["s1"], self.f["s2"] = self.st.empty(), self.st.empty()

    def _stored(self, t):
        if self.round_storage:
            t.copy_(t.float().double())

    def step(self):
        for form, o, s, c in _FORMS:
            self.st.advect_stage(form, self.f[o], self.f[s], self.f[c] if c else None, self.vel, self.dt)
            self._stored(self.f[o])

    def reinit(self):
        h = 0.5 * self.grid.dx[0]
        for form, o, s, c in _FORMS:
            self.st.reinit_stage(form, self.f[o], self.f[s], self.f[c] if c else None, None, False, h)
            self._stored(self.f[o])

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
    runs = {"ref64": _Run("double", n, False, False, None, device, velocity),
            "in64": _Run("double", n, True, False, None, device, velocity),
            "st64": _Run("double", n, True, True, None, device, velocity),
            "fp32": _Run("float", n, False, False, zero_tol, device, velocity)}
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
