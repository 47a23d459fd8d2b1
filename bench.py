"""Throughput of the WENO5 / TVD-RK3 level-set advection step on one MI355X.

Workload (default): 512^3 cells on the periodic unit cube, phi0 = signed
distance to the union of 64 random spheres (seed 5, radii U[0.02, 0.1]),
stored Taylor-Green velocity u = sin2pix cos2piy cos2piz,
v = -cos2pix sin2piy cos2piz, w = 0, CFL 0.5, 3-cell ghost layer.  One step
is three fused Runge-Kutta stage kernels (upwind HJ-WENO5 + advection RHS +
TVD-RK3 combination); periodic ghost images are written by the stages.

The FP32 step is timed (value), then the FP64 step on the same initial data
(fp64 block).  Kernel durations come from HIP events recorded around every
stage launch on the library's stream.  Algorithmic bytes per cell: stage 1
reads phi, u, v, w and writes phi1 (5 words); stages 2 and 3 also read phi^n
(6 words each): 17 words per cell-step.

cpu_baseline: the reference Fortran toolbox (flang build in oracle/_ref) or,
where it is absent, the C restatement, on one host core, advancing the same
workload at a smaller grid for a bounded time; all_cores: the same chain on
k-slabs from one host thread per usable core (harness-level parallelism).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "cell-updates/s per WENO5-RK3 step (FP32 vs FP64) + % HBM roofline; adjoint s"
HBM_PEAK_GBS = 8000.0
WORDS = {1: 5, 32: 6, 33: 6}   # algorithmic words per cell for each stage form
# FP32 leg: upwind switch at the double build's LSMLIB_ZERO_TOL (FP32 arithmetic
# throughout); the float build's 1e-5 drops the advection term wherever
# |V| < 1e-5, which dominates its error against FP64 on this workload.
SWITCH_TOL = 1e-11


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tiling", type=str, default="auto",
                    help="cfg,kz of the fused stage tiling, or 'auto' (tuned table)")
    ap.add_argument("--fp32-only", action="store_true", help="skip the FP64 leg (profiling runs)")
    ap.add_argument("--ghosts", choices=["synth", "images"], default="images",
                    help="synth: stages synthesise periodic ghost values; images: stages write and read "
                         "stored periodic ghost images")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-fields", action="store_true",
                    help="--only batched: build the initial patch pool on the host (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 2 GiB device-to-device copy (scale for HBM byte counters)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-patch and reinitialization measurements")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient (adjoint) measurement")
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference check of the gradient")
    ap.add_argument("--only", choices=["probe", "lf", "batched", "reinit", "adjoint", "config1", "config5", "throughput",
                                       "curvature", "curvature_step", "curvature_local", "curvature_local_step",
                                       "field_extension"],
                    default=None,
                    help="run one secondary measurement alone and print its JSON (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--velocity", choices=["taylor_green", "abc"], default="taylor_green",
                    help="velocity of the timed advection (abc: every component non-zero)")
    return ap.parse_args()


def run_gpu(real, n, steps, warmup, tiling, ghosts="synth", zero_tol=None, velocity="taylor_green"):
    import torch
    from lsmlib_amd import toolbox
    from lsmlib_amd.pylsmlib import RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, Stepper, stable_advection_dt
    from lsmlib_amd.workloads import line_elems, sphere_union_taylor_green
    from lsmlib_amd.tuning import choose_tiling

    dev = torch.device("cuda:0")
    dtype = torch.float32 if real == "float" else torch.float64
    grid, phi64, vel, bc = sphere_union_taylor_green(n, dtype=torch.float64, device=dev, align=line_elems(dtype),
                                                     velocity=velocity)
    phi = phi64.to(dtype).clone()
    del phi64
    dt = stable_advection_dt(vel, grid, 0.5, "double")  # one dt for both precisions
    vel = tuple(v.to(dtype).contiguous() for v in vel)
    cfg, kz = choose_tiling(real, n) if tiling == "auto" else tuple(int(x) for x in tiling.split(","))
    st = Stepper(grid, real, bc, ghosts=ghosts, zero_tol=zero_tol, tiling=(cfg, kz))
    st.fill_ghosts(phi)
    s1, s2 = st.empty(), st.empty()
    stream = torch.cuda.current_stream(dev)
    forms = (RK1, TVDRK3_STAGE2, TVDRK3_STAGE3)

    def step(events=None):
        for f in forms:
            out, ps, pc = {RK1: (s1, phi, None), TVDRK3_STAGE2: (s2, s1, phi),
                           TVDRK3_STAGE3: (phi, s2, phi)}[f]
            if events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            st.advect_stage(f, out, ps, pc, vel, dt)
            if events is not None:
                e1.record(stream)
                events.append((f, e0, e1))

    one_step = None
    for w in range(warmup):
        step()
        if w == 0:
            one_step = grid.interior(phi).clone()  # the field after exactly one step
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step(events)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = grid.interior_cells
    word = 4 if real == "float" else 8
    tot_bytes = sum(WORDS[f] * cells * word for f, _, _ in events)
    tot_ms = sum(e0.elapsed_time(e1) for _, e0, e1 in events)
    per_form = {}
    for f, e0, e1 in events:
        per_form.setdefault(f, []).append(e0.elapsed_time(e1))
    res = {
        "real": real, "grid": n, "dt": dt, "tiling": [cfg, kz],
        "value": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "kernel_ms_avg": tot_ms / len(events),
        "kernel_ms_by_stage": {str(f): sum(v) / len(v) for f, v in per_form.items()},
        "achieved_gbs": tot_bytes / (tot_ms * 1e-3) / 1e9,
        "algorithmic_bytes_per_launch": tot_bytes / len(events),
        "cells": cells,
        "steps_total": warmup + steps,
    }
    res["_one_step"] = one_step
    return res, grid, phi


def run_bandwidth_probe(mib=2048, reps=10, march_n=512):
    """Streaming HBM rate of memory-only kernels (lsm_amd_bandwidth_probe):
    copy, 5 reads + 1 write (the fused advection stage's stream count) with
    16-byte and with 4-byte accesses, and reads only, over `mib` MiB per
    stream, at several grid sizes; the best rate of each pattern is reported
    (GB/s = 1e9 B/s).  "march_pattern" is the stage kernels' own tile march
    with the arithmetic and the halo taken out (lsm_amd_march_probe, 32 x 8
    tiles, z-chunks of 16 planes, march_n^3 cells in the padded layout of the
    stage workloads), counted as 6 words per interior cell.  "valu_*" are
    FP32 vector-ALU issue rates (lsm_amd_valu_probe: eight independent chains
    per lane, no memory traffic), in lane-instructions/s: v_fma_f32, v_mul_f32
    and v_add_f32 with a scalar-register operand ("_sgpr"), with per-lane
    operands ("_vgpr", and the accumulating v_fmac_f32), with literal
    constants, the three-operand v_max3_f32 and the transcendental
    v_rcp_f32."""
    import ctypes
    import torch
    from lsmlib_amd import toolbox
    lib = toolbox("float")
    nbytes = mib << 20
    src = torch.ones(5 * nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    lib.set_stream(torch.cuda.current_stream().cuda_stream)
    out = {}
    for kind, name, moved in ((0, "copy", 2), (1, "read5_write1", 6), (3, "read5_write1_4byte", 6), (2, "read", 1)):
        best = None
        for wg in (1024, 2048, 4096, 8192, 16384, 32768, 65536):
            if kind == 2 and wg * 256 * 4 > nbytes:
                continue
            ms = lib.fused("lsm_amd_bandwidth_probe", kind, ctypes.c_void_p(src.data_ptr()),
                           ctypes.c_void_p(dst.data_ptr()), nbytes, wg, reps)
            gbs = moved * nbytes / (ms * 1e-3) / 1e9
            if best is None or gbs > best["gbs"]:
                best = {"gbs": gbs, "ms": ms, "workgroups": wg, "bytes_moved": moved * nbytes}
        out[name] = best
    del src, dst
    n = march_n
    px, py = n + 64, n + 6
    field = px * py * (n + 6)
    if 6 * field * 4 <= 6 * (mib << 20) + (256 << 20):
        src = torch.zeros(5 * field, dtype=torch.float32, device="cuda")
        dst = torch.empty(field, dtype=torch.float32, device="cuda")
        ms = lib.fused("lsm_amd_march_probe", ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                       n, px, py, 16, reps)
        out["march_pattern"] = {"gbs": 6 * 4 * n ** 3 / (ms * 1e-3) / 1e9, "ms": ms, "grid": n, "tile": [32, 8],
                                "kz": 16, "bytes_moved": 6 * 4 * n ** 3}
        del src, dst
    dst = torch.empty(16384 * 256, dtype=torch.float32, device="cuda")
    for op, name in ((0, "valu_fma_sgpr"), (1, "valu_rcp"), (2, "valu_mul_sgpr"), (3, "valu_add_sgpr"),
                     (4, "valu_fma_vgpr"), (5, "valu_fmac_vgpr"), (6, "valu_max3_vgpr"), (7, "valu_fma_literal")):
        best = None
        for wg in (2048, 4096, 16384):
            trips = (6 if op == 1 else 24) * 1024 * 1024 // wg
            ms = lib.fused("lsm_amd_valu_probe", ctypes.c_void_p(dst.data_ptr()), op, wg, trips, 5)
            rate = wg * 256 * trips * 32 / (ms * 1e-3)
            if best is None or rate > best["lane_instructions_per_s"]:
                best = {"lane_instructions_per_s": rate, "ms": ms, "workgroups": wg, "trips": trips}
        out[name] = best
    del dst
    torch.cuda.empty_cache()
    return out


def run_lf(n=512, steps=10, warmup=2):
    """WENO5 / TVD-RK3 advection with the local Lax-Friedrichs Hamiltonian on
    the headline workload (n^3 sphere union, Taylor-Green velocity, CFL 0.5),
    FP32 and FP64: per-stage kernel time from HIP events on the library
    stream, and the distance to the upwind step (the two agree where
    |v_a| >= zero_tol, so the difference is only the velocity's zero band)."""
    import torch
    from lsmlib_amd.pylsmlib import RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, Stepper, stable_advection_dt
    from lsmlib_amd.workloads import line_elems, sphere_union_taylor_green
    dev = torch.device("cuda:0")
    out = {}
    for real, dtype in (("float", torch.float32), ("double", torch.float64)):
        grid, phi64, vel, bc = sphere_union_taylor_green(n, dtype=torch.float64, device=dev, align=line_elems(dtype))
        dt = stable_advection_dt(vel, grid, 0.5, "double")
        phi = phi64.to(dtype).clone()
        del phi64
        vel = tuple(v.to(dtype).contiguous() for v in vel)
        st = Stepper(grid, real, bc, zero_tol=SWITCH_TOL if real == "float" else None)
        st.fill_ghosts(phi)
        up = phi.clone()
        s1, s2 = st.empty(), st.empty()
        for _ in range(warmup):
            st.advect_lf_rk3_step(phi, s1, s2, vel, dt)
            st.advect_rk3_step(up, s1, s2, vel, dt)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        ev = []
        for _ in range(steps):
            for f, o, ps, pc in ((RK1, s1, phi, None), (TVDRK3_STAGE2, s2, s1, phi), (TVDRK3_STAGE3, phi, s2, phi)):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                st.advect_lf_stage(f, o, ps, pc, vel, dt)
                e1.record(stream)
                ev.append((f, e0, e1))
        torch.cuda.synchronize()
        ms = sum(e0.elapsed_time(e1) for _, e0, e1 in ev) / len(ev)
        word = 4 if real == "float" else 8
        alg = sum(WORDS[f] for f, _, _ in ev) / len(ev) * grid.interior_cells * word
        for _ in range(steps):
            st.advect_rk3_step(up, s1, s2, vel, dt)
        torch.cuda.synchronize()
        diff = float((grid.interior(phi).double() - grid.interior(up).double()).abs().max()) / grid.dx[0]
        out[real] = {"kernel_ms_avg": ms, "cell_updates_per_s": grid.interior_cells / (3 * ms * 1e-3),
                     "achieved_gbs": alg / (ms * 1e-3) / 1e9, "steps": warmup + steps,
                     "max_diff_vs_upwind_over_dx": diff}
        print(f"# lax-friedrichs {real} {n}^3: stage {ms:.3f} ms, {out[real]['cell_updates_per_s']:.3e} "
              f"cell-updates/s, |LF - upwind| {diff:.2e} dx", file=sys.stderr, flush=True)
        del phi, up, s1, s2, vel, st
        torch.cuda.empty_cache()
    out["config"] = f"{n}^3 sphere union + Taylor-Green, local Lax-Friedrichs on plus/minus HJ-WENO5, TVD-RK3, CFL 0.5"
    return out


def run_batched(real="float", npatch=512, steps=20, reps=2, tiling="auto", host_fields=False):
    """512 uneven patches (extents U{16..96}, seed 3, dx 1/128, 3 ghosts,
    per-patch constant velocity), `steps` WENO5/TVD-RK3 steps captured into
    one hipGraph; replays are timed.  host_fields builds the initial pool on
    the host (profiling runs: no small set-up kernels)."""
    import torch
    from lsmlib_amd.patches import PatchLevel, random_patch_level
    from lsmlib_amd.workloads import line_elems
    desc = random_patch_level(npatch=npatch, lo=16, hi=96, seed=3, dx=1.0 / 128)
    dtype = torch.float32 if real == "float" else torch.float64
    tl = "auto" if tiling == "auto" else tuple(int(x) for x in tiling.split(","))
    lvl = PatchLevel(desc, real, align=line_elems(dtype), tiling=tl)
    phi = lvl.initial_phi(on_host=host_fields)
    s1, s2 = lvl.empty_pool(), lvl.empty_pool()
    dt = lvl.stable_dt(0.5)
    lvl.capture_rk3(steps, phi, s1, s2, dt)
    lvl.replay()  # warm-up
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        lvl.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    word = 4 if real == "float" else 8
    return {"real": real, "patches": npatch, "cells": lvl.cells, "steps_per_graph": steps,
            "graph_ms": ms, "value": lvl.cells * steps / (ms * 1e-3), "unit": "cell-updates/s",
            "stage_launches_per_graph": 3 * steps, "tiles": int(lvl.tiles), "workgroups": int(lvl.workgroups),
            "tiling": list(lvl.tiling), "ghosts": "signed-linear face ghosts synthesised in the stage kernel",
            "achieved_gbs": 8 * word * lvl.cells * steps / (ms * 1e-3) / 1e9,
            "words_per_cell_step": 8}


def run_reinit(n=256, iters=50):
    """Reinitialization, 256^3 on [-1,1]^3, phi0 = (r-0.5)(1+0.4 sin3x sin3y sin3z),
    smoothed sign from the frozen phi0 (use_phi0_for_sgn = 1), HJ-WENO5 +
    TVD-RK3, pseudo-dt 0.5 dx, 50 iterations; FP64 and FP32, FP32 error vs FP64."""
    import torch
    from lsmlib_amd.pylsmlib import Stepper
    from lsmlib_amd.workloads import line_elems, wavy_sphere_reinit
    dev = torch.device("cuda:0")
    out = {}
    res = {}
    for real, dtype in (("double", torch.float64), ("float", torch.float32)):
        grid, phi64, bc = wavy_sphere_reinit(n, torch.float64, dev, align=line_elems(dtype))
        # stored ghosts: the stages write their output's extrapolated face ghost layers
        st = Stepper(grid, real, bc, ghosts="images")
        phi = phi64.to(dtype).clone()  # a copy even when dtype is already float64
        phi0 = phi.clone()
        s1, s2 = st.empty(), st.empty()
        dt = 0.5 * grid.dx[0]
        st.fill_ghosts(phi)
        st.reinit_rk3_iteration(phi, s1, s2, dt, phi0=phi0, use_phi0=True)
        phi.copy_(phi0)
        st.fill_ghosts(phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            st.reinit_rk3_iteration(phi, s1, s2, dt, phi0=phi0, use_phi0=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[real] = (grid, phi)
        out[real] = {"ms_per_iteration": 1e3 * el / iters,
                     "value": grid.interior_cells * iters / el, "unit": "cell-updates/s"}
    err = fp32_error(*res["float"], *res["double"])
    out["fp32_error_vs_fp64"] = err
    out["config"] = (f"{n}^3 wavy sphere, use_phi0_for_sgn=1, {iters} TVD-RK3 iterations, dt=0.5dx, signed-linear "
                     "ghosts stored (each stage writes its output's face ghost layers)")
    return out


def run_adjoint(n=128, nsteps=40, every=8, reps=3, fd=True):
    """Gradient workload: 128^3 periodic [-1,1]^3, phi0 = |x| - 0.3, velocity
    sum_m P(k_m) c_m sin(pi k_m.x + theta_m) with 64 integer wavevectors
    (1 <= max|k_i| <= 4), c ~ N(0, 0.1^2) (seed 4); dt frozen at the CFL-0.5
    LSM3D_COMPUTE_STABLE_ADVECTION_DT of the initial velocity; 40 WENO5 /
    TVD-RK3 steps; loss sum (H_eps(phi_T) - H_eps(phi_tgt))^2, eps = 1.5 dx,
    phi_tgt advected with c + delta, delta ~ N(0, 0.02^2) (seed 9);
    gradients with respect to c (64 x 3) and phi0, checkpoint every 8 steps.

    adjoint_s: wall time of one forward + backward in FP32 (velocity from c,
    40 steps, loss, reverse sweep, chain to c), best of `reps`.  The FP64
    gradient is the reference for the FP32 one (upwind switch 1e-11 in both;
    fp32_grad_split says where the difference comes from); central differences of the
    FP64 loss check the FP64 gradient along 8 random unit directions in c,
    the 8 coefficients with the largest |dL/dc|, and 2 smooth phi0 bumps."""
    import numpy as np
    import torch
    from lsmlib_amd import toolbox
    from lsmlib_amd.adjoint import AdvectionRK3, advect_weno5_rk3, heaviside_mismatch
    from lsmlib_amd.workloads import adjoint_case
    dev = torch.device("cuda:0")
    grid, phi0, fv, _ = adjoint_case(n, torch.float64, dev)
    c0 = fv.c0.to(dev)
    v0 = fv.velocity(c0)
    dt = float(toolbox("double").lsm3dcomputestableadvectiondt_(v0[0].contiguous(), v0[1].contiguous(),
                                                               v0[2].contiguous(), grid.gb, grid.fb,
                                                               *grid.dx, 0.5))
    del v0
    rng = np.random.default_rng(9)
    delta = torch.tensor(rng.normal(0.0, 0.02, size=tuple(c0.shape)), dtype=torch.float64, device=dev)
    eps = 1.5 * grid.dx[0]
    out = {"config": f"{n}^3 periodic [-1,1]^3, sphere r=0.3, 64x3 Fourier modes |k|_inf 1..4, "
                     f"{nsteps} WENO5/TVD-RK3 steps, dt={dt:.6g} frozen, checkpoint every {every}, "
                     f"loss sum (H(phi_T)-H(phi_tgt))^2 eps=1.5dx"}
    grads, losses = {}, {}
    fd_dirs = []  # (name, c direction or None, phi0 direction or None, central difference of the FP64 loss)
    for real, dtype in (("double", torch.float64), ("float", torch.float32)):
        # the FP32 runner uses the double build's upwind switch (1e-11), so
        # both gradients see V_a-derivatives on the same cells
        runner = AdvectionRK3(grid, real, dev, checkpoint_every=every, zero_tol=1e-11 if real == "float" else None)
        with torch.no_grad():
            vt = tuple(v.to(dtype).contiguous() for v in fv.velocity(c0 + delta))
            target, _ = runner.run(phi0.to(dtype), vt, nsteps, dt, keep_checkpoints=False)
            del vt

        def loss_of(p0, c):
            vel = tuple(v.to(dtype).contiguous() for v in fv.velocity(c))
            phi = advect_weno5_rk3(p0, vel, runner, nsteps, dt)
            return heaviside_mismatch(phi, target, grid, eps)

        def fwd_bwd():
            c = c0.clone().requires_grad_(True)
            p0 = phi0.to(dtype).clone().requires_grad_(True)
            L = loss_of(p0, c)
            L.backward()
            return L, c.grad, p0.grad

        fwd_bwd()
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            L, gc, gp = fwd_bwd()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        grads[real] = (gc.double(), gp.double())
        losses[real] = float(L.detach())
        out[real] = {"forward_backward_s": best, "loss": float(L.detach())}
        if real == "double" and fd:
            checks = []
            with torch.no_grad():
                def fd_dir(dc=None, dp=None, h=1e-4):
                    cp = c0 + h * dc if dc is not None else c0
                    cm = c0 - h * dc if dc is not None else c0
                    pp = phi0 + h * dp if dp is not None else phi0
                    pm = phi0 - h * dp if dp is not None else phi0
                    return (float(loss_of(pp, cp)) - float(loss_of(pm, cm))) / (2 * h)
                for _ in range(8):
                    e = torch.tensor(rng.normal(size=tuple(c0.shape)), dtype=torch.float64, device=dev)
                    e /= e.norm()
                    ad, fdv = float((gc * e).sum()), fd_dir(dc=e)
                    checks.append(("c_random", abs(fdv - ad) / abs(ad)))
                    fd_dirs.append(("c_random", e, None, fdv))
                top = torch.argsort(gc.abs().flatten(), descending=True)[:8]
                for idx in top.tolist():
                    e = torch.zeros_like(c0).flatten()
                    e[idx] = 1.0
                    e = e.reshape(c0.shape)
                    ad, fdv = float(gc.flatten()[idx]), fd_dir(dc=e)
                    checks.append((f"c[{idx // 3},{idx % 3}]", abs(fdv - ad) / abs(ad)))
                    fd_dirs.append((f"c[{idx // 3},{idx % 3}]", e, None, fdv))
                X, Y, Z = grid.centers(torch.float64, dev)
                for w in ((1.0, 2.0, 0.5), (3.0, 1.0, 2.0)):
                    e = torch.sin(np.pi * (w[0] * X + w[1] * Y) + w[2] * Z) * torch.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 0.2)
                    ad, fdv = float((gp * e).sum()), fd_dir(dp=e)
                    checks.append(("phi0_bump", abs(fdv - ad) / abs(ad)))
                    fd_dirs.append(("phi0_bump", None, e, fdv))
                del X, Y, Z
            out["fd_check_fp64"] = {"max_rel_err": max(v for _, v in checks), "n_directions": len(checks),
                                    "h": 1e-4, "per_direction": {f"{k}#{q}": v for q, (k, v) in enumerate(checks)}}
        del runner, target
        torch.cuda.empty_cache()
    if fd_dirs:
        # the FP32 gradient (FP32 forward, checkpoints and adjoint) along the
        # same directions against the same central differences of the FP64 loss
        g32c, g32p = grads["float"]
        errs = [(k, abs(fdv - float((g32c * dc).sum() if dc is not None else (g32p * dp).sum())) / abs(fdv))
                for k, dc, dp, fdv in fd_dirs]
        out["fd_check_fp32"] = {"max_rel_err": max(v for _, v in errs), "n_directions": len(errs),
                                "reference": "central differences of the FP64 loss (h 1e-4)",
                                "per_direction": {f"{k}#{q}": v for q, (k, v) in enumerate(errs)}}
    out["fp32_grad_rel_err_vs_fp64"] = {
        "c": float((grads["float"][0] - grads["double"][0]).norm() / grads["double"][0].norm()),
        "phi0": float((grads["float"][1] - grads["double"][1]).norm() / grads["double"][1].norm())}
    from lsmlib_amd.precision import gradient_precision_split
    out["fp32_grad_split"] = gradient_precision_split(n, nsteps, every, dev)
    out["adjoint_s"] = out["float"]["forward_backward_s"]
    out["dt"] = dt
    return out


def run_config1(n=128, steps=100):
    """Plain FP64 advection, 128^3 on [-1,1]^3, sphere r=0.35 at (0.35,0,0), rigid
    rotation u=(-y,x,0), signed-linear ghosts, HJ-WENO5 + TVD-RK3, CFL 0.5,
    100 steps: the GPU FP64 fused path against the reference Fortran FP64
    toolbox (flang build) run on this host over the same steps."""
    import numpy as np
    import torch
    from lsmlib_amd.pylsmlib import Stepper, stable_advection_dt
    from lsmlib_amd.workloads import rotating_sphere
    from oracle.oracle import oracle_lib, ref_lib
    from oracle.steps import OracleStepper, fill_ghosts
    dev = torch.device("cuda:0")
    grid, phi, vel, bc = rotating_sphere(n, torch.float64, dev)
    st = Stepper(grid, "double", bc)
    dt = stable_advection_dt(vel, grid, 0.5)
    p = phi.clone()
    s1, s2 = st.empty(), st.empty()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st.advect_rk3_step(p, s1, s2, vel, dt)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    got = grid.interior(p).cpu().numpy()
    lib = ref_lib("double")
    kind = "reference"
    if lib is None:
        lib, kind = oracle_lib("double"), "port"
    ph = phi.cpu().numpy().copy()
    vn = tuple(v.cpu().numpy().copy() for v in vel)
    op = OracleStepper(lib, grid, bc)
    fill_ghosts(lib, ph, grid, bc)
    t0 = time.perf_counter()
    for _ in range(steps):
        op.advect_rk3_step(ph, vn, dt)
    cpu_s = time.perf_counter() - t0
    want = grid.interior(ph)
    cells = grid.interior_cells * steps
    return {"config": f"{n}^3 rotating sphere, signed-linear ghosts, CFL 0.5, {steps} steps, FP64",
            "gpu_s": gpu_s, "gpu_value": cells / gpu_s, "cpu_s": cpu_s, "cpu_value": cells / cpu_s,
            "cpu_kind": kind, "cpu_cores": 1, "unit": "cell-updates/s",
            "max_rel_diff_vs_reference": float(np.abs(got - want).max() / np.abs(want).max())}


def run_config5(n=512, steps=200, reinit_every=10, reinit_iters=5):
    """Config 5 in full: 512^3 periodic sphere union (64 spheres, seed 5),
    Taylor-Green velocity, WENO5/TVD-RK3 advection with 5 reinitialization
    iterations (sign from the current phi, pseudo-dt 0.5 dx) every 10 steps,
    200 steps, FP32 (upwind switch 1e-11) and FP64; error of FP32 vs FP64."""
    import torch
    from lsmlib_amd.pylsmlib import Stepper, stable_advection_dt
    from lsmlib_amd.workloads import line_elems, sphere_union_taylor_green
    dev = torch.device("cuda:0")
    out, res = {}, {}
    for real, dtype in (("float", torch.float32), ("double", torch.float64)):
        grid, phi64, vel, bc = sphere_union_taylor_green(n, dtype=torch.float64, device=dev,
                                                         align=line_elems(dtype))
        dt = stable_advection_dt(vel, grid, 0.5, "double")
        phi = phi64.to(dtype).clone()
        del phi64
        vel = tuple(v.to(dtype).contiguous() for v in vel)
        st = Stepper(grid, real, bc, zero_tol=SWITCH_TOL if real == "float" else None)
        st.fill_ghosts(phi)
        s1, s2 = st.empty(), st.empty()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for step in range(steps):
            st.advect_rk3_step(phi, s1, s2, vel, dt)
            if (step + 1) % reinit_every == 0:
                for _ in range(reinit_iters):
                    st.reinit_rk3_iteration(phi, s1, s2, 0.5 * grid.dx[0])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[real] = {"seconds": el, "value": grid.interior_cells * steps / el, "unit": "cell-updates/s"}
        res[real] = (grid, phi)
        del s1, s2, vel
        torch.cuda.empty_cache()
    out["fp32_error_vs_fp64"] = fp32_error(*res["float"], *res["double"])
    del res
    torch.cuda.empty_cache()
    from lsmlib_amd.precision import precision_split
    out["fp32_error_split"] = precision_split(n, steps, reinit_every, reinit_iters)
    out["config"] = (f"{n}^3 periodic sphere union + Taylor-Green, {steps} WENO5/TVD-RK3 steps with "
                     f"{reinit_iters} reinitialization iterations every {reinit_every} steps")
    return out


def run_throughput(n=1024, steps=5, warmup=1):
    """FP32 advection throughput at 1024^3 (same workload as the headline)."""
    r, _, _ = run_gpu("float", n, steps, warmup, "auto", "images", zero_tol=SWITCH_TOL)
    return {"grid": n, "value": r["value"], "unit": "cell-updates/s", "ms_per_step": r["ms_per_step"],
            "kernel_ms_avg": r["kernel_ms_avg"], "achieved_gbs": r["achieved_gbs"], "tiling": r["tiling"]}


CURV_PUBLISHED_S = {0.04: 105.0, 0.02: 1276.0}
# t_local of the narrow-band example, curvature_example/README:143-144
LOCAL_PUBLISHED_S = {0.04: 102.0, 0.02: 1086.0}   # curvature_example/README:141-144 (regular method)


def run_curvature(dxs=(0.04, 0.02), big_dx=1.5 / 512, big_steps=10, cpu=True, runs=True):
    """The curvature example (phi_t + a|grad phi| = b kappa|grad phi|, a = 0.4,
    b = 0.05, mask = three unit spheres; curvature_model3d.c:41-626) on its
    default problem.  (1) Whole runs to the example's own stopping rule at
    dx 0.04 and 0.02, FP64 and FP32, timed end to end (host loop included)
    beside the published timings; at dx 0.04 the reference Fortran toolbox
    chain (oracle/_ref) runs the same loop on one host core and the FP64
    fields are compared.  (2) The model step on the same geometry at
    dx = 1.5/512 (512 x 512 x 341 cells): per-kernel HIP-event times of the
    dt reduction, the two fused stages and the ghost fills."""
    import numpy as np
    import torch
    from lsmlib_amd.curvature import CurvatureFlow
    from lsmlib_amd.pylsmlib import RK1, TVDRK2_STAGE2
    from lsmlib_amd.workloads import pore_throat

    dev = torch.device("cuda:0")
    out = {"model": "a=0.4, b=0.05, cfl 0.5, ENO2 / TVD-RK2, mask, reinit every 0.1 (11 iterations)",
           "runs": {}}
    finals = {}
    if not runs:
        dxs, cpu = (), False
    for real, dt_ in (("double", torch.float64), ("float", torch.float32)):
        for dx in dxs:
            g, phi, mask = pore_throat(dx, dt_, dev)
            m = CurvatureFlow(g, real, mask, device=dev)
            p = phi.clone()
            m.run(p, max_outer=1)          # warm-up (kernel loads, allocations)
            p = phi.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = m.run(p)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            out["runs"][f"{real}_dx{dx}"] = {
                "grid": list(g.n), "seconds": el, "steps": st["steps"], "intervals": st["outer"],
                "reinit_iterations": st["reinit_iters"], "t_final": st["t"], "vol_phi": st["vol_phi"],
                "max_change_last_interval": st["max_abs_err"],
                "published_reference_s": CURV_PUBLISHED_S.get(dx)}
            finals[(real, dx)] = (g, p)
            print(f"# curvature {real} dx={dx} {g.n}: {el:.3f} s, {st['steps']} steps", file=sys.stderr, flush=True)
    # FP32 against FP64 after the whole run, near the interface
    for dx in dxs:
        g, p64 = finals[("double", dx)]
        _, p32 = finals[("float", dx)]
        e = interior_error(p32, p64, g)
        out["runs"][f"float_dx{dx}"]["error_vs_fp64"] = e
    if cpu:
        from oracle.oracle import oracle_lib, ref_lib
        from oracle.steps import CurvatureFlowOracle
        lib = ref_lib("double")
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib("double"), "port"
        g, phi, mask = pore_throat(0.04)
        pc = phi.numpy().copy()
        t0 = time.perf_counter()
        stc = CurvatureFlowOracle(lib, g, mask.numpy().copy()).run(pc)
        el = time.perf_counter() - t0
        gpu_p = finals[("double", 0.04)][1].cpu().numpy()
        out["cpu_dx0.04"] = {"kind": kind, "cores": 1, "seconds": el, "steps": stc["steps"],
                             "identical_final_field_fp64": bool(np.array_equal(gpu_p, pc)),
                             "max_abs_diff_fp64": float(np.abs(gpu_p - pc).max())}
        print(f"# curvature cpu ({kind}) dx=0.04: {el:.2f} s, same field: "
              f"{out['cpu_dx0.04']['identical_final_field_fp64']}", file=sys.stderr, flush=True)
    # throughput of the model step at 512 x 512 x 341
    big = {}
    for real, dt_ in (("float", torch.float32), ("double", torch.float64)):
        g, phi, mask = pore_throat(big_dx, dt_, dev)
        m = CurvatureFlow(g, real, mask, device=dev)
        p = phi.clone()
        m.fill_ghosts(p, masked=False)
        esz = 4 if real == "float" else 8
        cells = g.n[0] * g.n[1] * g.n[2]
        ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
        names = ("dt", "stage1", "ghosts1", "stage2", "ghosts2")
        acc = {k: 0.0 for k in names}
        m.step(p, None, 0.0, 1.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(big_steps):
            e = [ev() for _ in range(6)]
            e[0].record()
            dt = m._choose_dt(m.normal_dt(p), 0.0, 1.0)
            e[1].record()
            m.curvature_stage(RK1, m.s1, p, None, dt)
            e[2].record()
            m.fill_ghosts(m.s1)
            e[3].record()
            m.curvature_stage(TVDRK2_STAGE2, p, m.s1, p, dt)
            e[4].record()
            m.fill_ghosts(p)
            e[5].record()
            e[5].synchronize()
            for q, k in enumerate(names):
                acc[k] += e[q].elapsed_time(e[q + 1])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = {k: v / big_steps for k, v in acc.items()}
        bytes_cell = {"stage1": 3 * esz, "stage2": 4 * esz, "dt": esz}
        big[real] = {"grid": list(g.n), "cells": cells, "steps": big_steps, "ms_per_step": 1e3 * el / big_steps,
                     "cell_steps_per_s": cells * big_steps / el, "kernel_ms": ms,
                     "stage2_gbs": bytes_cell["stage2"] * cells / ms["stage2"] / 1e6,
                     "stage1_gbs": bytes_cell["stage1"] * cells / ms["stage1"] / 1e6,
                     "dt_gbs": bytes_cell["dt"] * cells / ms["dt"] / 1e6,
                     "algorithmic_bytes_per_cell": bytes_cell}
        print(f"# curvature step {real} {g.n}: {big[real]['ms_per_step']:.3f} ms/step, stage2 "
              f"{ms['stage2']:.3f} ms ({big[real]['stage2_gbs']:.0f} GB/s)", file=sys.stderr, flush=True)
        del m, p, phi, mask
        torch.cuda.empty_cache()
    out["step_512x512x341"] = big
    return out


def run_curvature_local(dxs=(0.04, 0.02), cpu=True, cpu_intervals=3, big_dx=1.5 / 512, big_steps=5):
    """The curvature example's narrow-band loop (curvatureModelMedium3dLocalMainLoop,
    curvature_model3d_local.c:52-728) on its default problem: (1) whole runs to
    the example's stopping rule at dx 0.04 and 0.02, FP64 and FP32, beside the
    published t_local (curvature_example/README:135-145); (2) at dx 0.04 the
    reference Fortran narrow-band routines (oracle/_ref) run the same loop on
    one host core for the first `cpu_intervals` intervals and the FP64 fields
    are compared with the GPU's after the same intervals; (3) the localized
    step at dx = 1.5/512 (512 x 512 x 341 cells), split by HIP events into the
    band build, the two stages and the reinitialization check."""
    import numpy as np
    import torch
    from lsmlib_amd.curvature_local import LocalCurvatureFlow
    from lsmlib_amd.workloads import pore_throat

    dev = torch.device("cuda:0")
    out = {"model": "a=0.4, b=0.05, cfl 0.5, ENO2 / TVD-RK2, mask; band gamma 4dx, beta 2dx, levels 0-3",
           "published_t_local_s": LOCAL_PUBLISHED_S, "runs": {}}
    for real, dt_ in (("double", torch.float64), ("float", torch.float32)):
        for dx in dxs:
            g, phi, mask = pore_throat(dx, dt_, dev)
            m = LocalCurvatureFlow(g, real, mask, device=dev)
            p = phi.clone()
            m.run(p, max_outer=1)
            p = phi.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = m.run(p)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            out["runs"][f"{real}_dx{dx}"] = {
                "grid": list(g.n), "seconds": el, "steps": st["steps"], "intervals": st["outer"],
                "reinitializations": st["reinits"], "triggers": st["triggers"], "t_final": st["t"],
                "vol_phi": st["vol_phi"], "mean_band_fraction": float(np.mean(st["band_fraction"])),
                "ms_per_step": 1e3 * el / max(st["steps"], 1)}
            print(f"# curvature local {real} dx={dx} {g.n}: {el:.3f} s, {st['steps']} steps, band fraction "
                  f"{np.mean(st['band_fraction']):.3f}", file=sys.stderr, flush=True)
    if cpu:
        from oracle.oracle import oracle_lib, ref_lib
        from oracle.steps import LocalCurvatureFlowOracle
        lib = ref_lib("double")
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib("double"), "port"
        g, phi, mask = pore_throat(0.04)
        pc = phi.numpy().copy()
        t0 = time.perf_counter()
        stc = LocalCurvatureFlowOracle(lib, g, mask.numpy().copy(), dt_lib=oracle_lib("double")).run(
            pc, max_outer=cpu_intervals)
        el = time.perf_counter() - t0
        gg, pd, md = pore_throat(0.04, torch.float64, dev)
        LocalCurvatureFlow(gg, "double", md, device=dev).run(pd, max_outer=cpu_intervals)
        torch.cuda.synchronize()
        same = bool(np.array_equal(pd.cpu().numpy(), pc))
        out["cpu_dx0.04"] = {"kind": kind, "cores": 1, "seconds": el, "steps": stc["steps"],
                             "sample": f"first {cpu_intervals} intervals (t = 0 .. {0.1 * cpu_intervals:g})",
                             "identical_field_fp64": same}
        print(f"# curvature local cpu ({kind}) dx=0.04, {cpu_intervals} intervals: {el:.2f} s, same field: {same}",
              file=sys.stderr, flush=True)
    big = {}
    for real, dt_ in (("float", torch.float32), ("double", torch.float64)):
        g, phi, mask = pore_throat(big_dx, dt_, dev)
        m = LocalCurvatureFlow(g, real, mask, device=dev)
        p = phi.clone()
        m.box.fill_ghosts(p, masked=False)
        m.reinitialize_box(p, m.gamma + g.dx[0])
        ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
        rec = {}
        # the library's step (band build + fused stages), then the same step
        # sequenced routine by routine from Python, each after a warm-up step
        for name, step, check in (("native", m.step_native, m.needs_reinit_native),
                                  ("composed", m.step, m.needs_reinit)):
            step(p, 0.0, 1.0)
            check(p, 0.0, 0.0, 0.0)
            torch.cuda.synchronize()
            acc = {"band": 0.0, "step": 0.0, "check": 0.0}
            for _ in range(big_steps):
                e = [ev() for _ in range(4)]
                e[0].record()
                m.band(p)          # the band build alone (the step builds it again, identically)
                e[1].record()
                step(p, 0.0, 1.0)
                e[2].record()
                check(p, 0.0, 0.0, 0.0)
                e[3].record()
                e[3].synchronize()
                acc["band"] += e[0].elapsed_time(e[1])
                acc["step"] += e[1].elapsed_time(e[2])
                acc["check"] += e[2].elapsed_time(e[3])
            rec[name] = {k: v / big_steps for k, v in acc.items()}
        torch.cuda.synchronize()
        cells = g.n[0] * g.n[1] * g.n[2]
        frac = float(m.n_hi[2] - m.n_lo[0] + 1) / m.n
        step_ms = rec["native"]["step"] + rec["native"]["check"]
        big[real] = {"grid": list(g.n), "band_points": int(m.n_hi[3] - m.n_lo[0] + 1),
                     "band_fraction_levels_0_2": frac, "ms_band_build": rec["native"]["band"],
                     "ms_per_step": step_ms, "ms_step_only": rec["native"]["step"],
                     "ms_reinit_check": rec["native"]["check"],
                     "ms_per_step_composed": rec["composed"]["step"] + rec["composed"]["check"],
                     "cell_steps_per_s": cells / (step_ms * 1e-3), "steps": big_steps}
        print(f"# curvature local step {real} {g.n}: {step_ms:.3f} ms/step with the reinitialization check "
              f"(band build {rec['native']['band']:.3f} ms, fraction {frac:.3f}); routine by routine "
              f"{big[real]['ms_per_step_composed']:.3f} ms", file=sys.stderr, flush=True)
        del m, p, phi, mask
        torch.cuda.empty_cache()
    out["step_512x512x341"] = big
    return out


def run_field_extension(n=512, iters=10):
    """Field extension (FieldExtensionAlgorithm.cc:195-440) on n^3: phi =
    distance to the union of 64 spheres, S = sin(2 pi x) y + z; the PHI_UPWIND
    signed normal and the extension velocity once, then TVD-RK3 iterations
    (fused WENO5 advection stages with that velocity).  Times the normal
    computation and the iterations with HIP events, FP32 and FP64."""
    import torch
    from lsmlib_amd.field_extension import FieldExtension
    from lsmlib_amd.pylsmlib import BC_LINEAR
    from lsmlib_amd.workloads import line_elems, sphere_union_taylor_green

    out = {}
    for real, dt_ in (("float", torch.float32), ("double", torch.float64)):
        grid, phi, vel, _ = sphere_union_taylor_green(n, dtype=dt_, device="cuda", align=line_elems(dt_))
        del vel
        X, Y, Z = grid.centers(dt_, "cuda")
        S = (torch.sin(2 * torch.pi * X) * Y + Z).contiguous()
        del X, Y, Z
        fe = FieldExtension(grid, real, (BC_LINEAR,) * 6, device="cuda")
        fe.stepper.fill_ghosts(phi)
        fe.normals(phi)
        fe.step(S)
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        fe.normals(phi)
        e[1].record()
        for _ in range(iters):
            fe.step(S)
        e[2].record()
        e[2].synchronize()
        ms_it = e[1].elapsed_time(e[2]) / iters
        cells = grid.interior_cells
        out[real] = {"grid": [n, n, n], "normals_ms": e[0].elapsed_time(e[1]), "ms_per_iteration": ms_it,
                     "cell_iterations_per_s": cells / (ms_it * 1e-3), "iterations": iters}
        print(f"# field extension {real} {n}^3: normals {out[real]['normals_ms']:.2f} ms, "
              f"{ms_it:.3f} ms/iteration", file=sys.stderr, flush=True)
        del fe, phi, S
        torch.cuda.empty_cache()
    return out


def fp32_error(grid32, phi32, grid64, phi64):
    """Max |phi32 - phi64| over the fill box and near the interface (|phi64| < 3 dx)."""
    return interior_error(grid32.interior(phi32), grid64.interior(phi64), grid64)


def interior_error(a32, b64, grid64):
    """Max |a32 - b64| globally and near the interface (|b64| < 3 dx); near the
    interface also the fraction of cells above 1e-5 dx and the 99.99th
    percentile (the maximum sits at the few cells where the union of spheres
    has a kink)."""
    import torch
    a = a32.to(torch.float64)
    b = b64
    err = (a - b).abs()
    h = min(grid64.dx)
    near = b.abs() < 3 * h
    out = {"max_abs_global": float(err.max()),
           "max_abs_near_interface": float(err[near].max()) if bool(near.any()) else None,
           "max_near_interface_over_dx": (float(err[near].max()) / h) if bool(near.any()) else None}
    if bool(near.any()):
        e = err[near] / h
        out["near_cells"] = int(e.numel())
        out["near_cells_above_1e-5_dx"] = int((e > 1e-5).sum())
        sample = e if e.numel() <= 4_000_000 else e[torch.randperm(e.numel(), device=e.device)[:4_000_000]]
        out["near_p9999_over_dx"] = float(torch.quantile(sample, 0.9999))
    return out


def cpu_baseline(seconds):
    """Reference toolbox chain on one host core, FP32 build, 48^3 sample."""
    import numpy as np
    import torch
    from oracle.oracle import oracle_lib, ref_lib
    from oracle.steps import OracleStepper, fill_ghosts
    from lsmlib_amd.workloads import sphere_union_taylor_green

    lib = ref_lib("float")
    kind = "reference"
    if lib is None:
        lib, kind = oracle_lib("float"), "port"
    n = 48
    grid, phi, vel, bc = sphere_union_taylor_green(n, dtype=torch.float32, device="cpu")
    ph = phi.numpy().copy()
    vn = tuple(v.numpy().copy() for v in vel)
    op = OracleStepper(lib, grid, bc)
    fill_ghosts(lib, ph, grid, bc)
    dt = lib.lsm3dcomputestableadvectiondt_(vn[0], vn[1], vn[2], grid.gb, grid.fb, *grid.dx, 0.5)
    steps = 0
    t0 = time.perf_counter()
    while True:
        op.advect_rk3_step(ph, vn, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    out = {"value": grid.interior_cells * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": kind,
           "sample": f"{n}^3 periodic sphere-union Taylor-Green advection, FP32 build, WENO5/TVD-RK3 "
                     f"toolbox chain, {steps} steps in {el:.1f} s"}
    out["all_cores"] = cpu_baseline_slabs(lib, kind, max(4.0, seconds / 2))
    return out


def host_cpu():
    """(CPU model, usable cores) of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline_slabs(lib, kind, seconds, n=128):
    """The same toolbox chain with every routine called on k-slabs of the fill
    box from one host thread per usable core (at most OMP_NUM_THREADS, at most
    16): the harness-level parallelism of the serial reference.  One step is
    first checked against the serial chain bit for bit."""
    import numpy as np
    import torch
    from oracle.steps import OracleStepper, SlabOracleStepper, fill_ghosts
    from lsmlib_amd.workloads import sphere_union_taylor_green

    model, cores = host_cpu()
    threads = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)), 16))
    grid, phi, vel, bc = sphere_union_taylor_green(n, dtype=torch.float32, device="cpu")
    ph = phi.numpy().copy()
    vn = tuple(v.numpy().copy() for v in vel)
    fill_ghosts(lib, ph, grid, bc)
    dt = lib.lsm3dcomputestableadvectiondt_(vn[0], vn[1], vn[2], grid.gb, grid.fb, *grid.dx, 0.5)
    ref = ph.copy()
    OracleStepper(lib, grid, bc).advect_rk3_step(ref, vn, dt)
    op = SlabOracleStepper(lib, grid, bc, threads)
    try:
        op.advect_rk3_step(ph, vn, dt)
        same = bool(np.array_equal(ph, ref))
        steps = 0
        t0 = time.perf_counter()
        while True:
            op.advect_rk3_step(ph, vn, dt)
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
    finally:
        op.close()
    return {"value": grid.interior_cells * steps / el, "unit": "cell-updates/s", "cores": op.threads,
            "kind": kind + ", harness-parallel", "cpu_model": model, "usable_cores": cores,
            "identical_to_serial": same,
            "sample": f"{n}^3, the same chain on {op.threads} k-slabs per routine, {steps} steps in {el:.1f} s"}


def read_traffic():
    p = REPO / "profiles" / "traffic_advect_stage_512_fp32.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


# FP32 vector-ALU peak of MI355X (157.3 TFLOP/s of FMA): 256 CUs x 128 lane-instructions
# per clock x 2.4 GHz.  The library is compiled without packed FP32; plain FP32
# instructions with per-lane or constant operands still issue near this rate on gfx950
# (lsm_amd_valu_probe), those with a scalar-register operand, three-operand min/max and
# transcendentals at half of it or less.
VALU_PEAK_LANE_INSTR = 256 * 128 * 2.4e9


def valu_issue_peak(probe):
    """The best FP32 issue rate of the probe's instruction chains,
    lane-instructions/s."""
    rates = [probe[k]["lane_instructions_per_s"] for k in probe if k.startswith("valu_")]
    return max(rates) if rates else None


def valu_roofline(r32, measured=None):
    """The stage kernel against the VALU issue rate: measured instructions per
    cell x cells per launch / the live average launch time, over the nominal
    rate and over the rate the FMA probe reached in the same run."""
    per_cell = read_valu()
    if per_cell is None or not r32.get("kernel_ms_avg"):
        return None
    achieved = per_cell * r32["cells"] / (r32["kernel_ms_avg"] * 1e-3)
    return {"lane_instructions_per_cell": per_cell, "achieved": achieved, "peak": VALU_PEAK_LANE_INSTR,
            "unit": "lane-instructions/s", "frac": achieved / VALU_PEAK_LANE_INSTR,
            "measured_peak": measured, "frac_of_measured": achieved / measured if measured else None}


def read_valu():
    """VALU lane-instructions per cell of the FP32 stage kernel (SQ_INSTS_VALU
    x 64 / cells, profiles/valu_advect_stage_512_fp32.json)."""
    p = REPO / "profiles" / "valu_advect_stage_512_fp32.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("valu_lane_instructions_per_cell")
        except Exception:
            return None
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the library runs on a "
              "single GPU (sharding into z-slabs with halo exchange is designed, not implemented)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a visible GPU", file=sys.stderr)
        sys.exit(1)
    if a.copy_calibration:
        n = (2 << 30) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                          "launches": 3}), flush=True)
        return

    if a.only == "lf":
        print(json.dumps({"lax_friedrichs_512": run_lf(a.n)}), flush=True)
        return
    if a.only == "probe":
        print(json.dumps({"bandwidth_probe": run_bandwidth_probe()}), flush=True)
        return
    if a.only == "adjoint":
        print(json.dumps({"adjoint": run_adjoint(fd=not a.no_fd)}), flush=True)
        return
    if a.only == "batched":
        print(json.dumps({"batched_patches_fp32": run_batched("float", tiling=a.tiling, host_fields=a.host_fields)}),
              flush=True)
        return
    if a.only == "reinit":
        print(json.dumps({"reinit_256": run_reinit()}), flush=True)
        return
    if a.only == "config1":
        print(json.dumps({"config1_fp64_128": run_config1()}), flush=True)
        return
    if a.only == "config5":
        print(json.dumps({"config5_512_full": run_config5()}), flush=True)
        return
    if a.only == "throughput":
        print(json.dumps({"throughput_1024_fp32": run_throughput()}), flush=True)
        return
    if a.only == "curvature":
        print(json.dumps({"curvature_example": run_curvature(cpu=not a.no_cpu_baseline)}), flush=True)
        return
    if a.only == "curvature_local_step":  # the localized 512 x 512 x 341 step alone (profiling runs)
        print(json.dumps({"curvature_local_example": run_curvature_local(dxs=(), cpu=False, big_steps=a.steps)}),
              flush=True)
        return
    if a.only == "curvature_local":
        print(json.dumps({"curvature_local_example": run_curvature_local(cpu=not a.no_cpu_baseline)}), flush=True)
        return
    if a.only == "field_extension":
        print(json.dumps({"field_extension_512": run_field_extension()}), flush=True)
        return
    if a.only == "curvature_step":  # the 512 x 512 x 341 model step alone (profiling runs)
        print(json.dumps({"curvature_example": run_curvature(runs=False, big_steps=a.steps)}), flush=True)
        return
    r32, grid32, phi32 = run_gpu("float", a.n, a.steps, a.warmup, a.tiling, a.ghosts, zero_tol=SWITCH_TOL,
                                 velocity=a.velocity)
    print(f"# fp32: {r32['value']:.4e} cell-updates/s, {r32['ms_per_step']:.3f} ms/step, "
          f"kernel {r32['kernel_ms_avg']:.3f} ms, {r32['achieved_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    r64 = None
    err = err_build_tol = err_one = None
    if not a.fp32_only:
        r64, grid64, phi64 = run_gpu("double", a.n, a.steps, a.warmup, a.tiling, a.ghosts, velocity=a.velocity)
        print(f"# fp64: {r64['value']:.4e} cell-updates/s, {r64['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
        err = fp32_error(grid32, phi32, grid64, phi64)
        err["after_steps"] = r32["steps_total"]
        err_one = interior_error(r32.pop("_one_step"), r64.pop("_one_step"), grid64)
        del phi32
        # the same FP32 run with the float build's LSMLIB_ZERO_TOL (1e-5) in the upwind switch
        _, _, phi32b = run_gpu("float", a.n, a.steps, a.warmup, a.tiling, a.ghosts, zero_tol=None,
                               velocity=a.velocity)
        err_build_tol = fp32_error(grid32, phi32b, grid64, phi64)
        err_build_tol["after_steps"] = r32["steps_total"]
        print(f"# fp32 error vs fp64 near interface: max {err['max_near_interface_over_dx']:.3e} dx "
              f"(switch 1e-11; {err_build_tol['max_near_interface_over_dx']:.3e} dx with 1e-5), "
              f"{err['near_cells_above_1e-5_dx']} of {err['near_cells']} cells above 1e-5 dx",
              file=sys.stderr, flush=True)
        del phi64, phi32b
    phi32 = None
    torch.cuda.empty_cache()
    split = None
    if not a.fp32_only:
        from lsmlib_amd.precision import precision_split
        split = precision_split(a.n, a.warmup + a.steps, velocity=a.velocity)
        print(f"# fp32 error split after {a.warmup + a.steps} steps (max near interface / dx): fp32 "
              f"{split['fp32_vs_fp64']['max_over_dx']:.2e}, inputs only {split['inputs_only']['max_over_dx']:.2e}, "
              f"fp32 storage only {split['storage_only']['max_over_dx']:.2e} (phi alone "
              f"{split['storage_phi_only']['max_over_dx']:.2e}, s1 and s2 alone "
              f"{split['storage_stages_only']['max_over_dx']:.2e}), one step "
              f"{split['one_step']['max_over_dx']:.2e}", file=sys.stderr, flush=True)
    probe = run_bandwidth_probe()
    best5 = max(probe["read5_write1"]["gbs"], probe["read5_write1_4byte"]["gbs"])
    print(f"# measured streaming rates: copy {probe['copy']['gbs']:.0f} GB/s, 5 reads + 1 write "
          f"{best5:.0f} GB/s, reads {probe['read']['gbs']:.0f} GB/s, the stage's tile march without arithmetic "
          f"{probe.get('march_pattern', {}).get('gbs', float('nan')):.0f} GB/s; FP32 VALU issue "
          f"{(valu_issue_peak(probe) or float('nan')) / 1e12:.1f} T lane-instructions/s (best plain ALU chain)",
          file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a.cpu_seconds)
        print(f"# cpu: {cpu['value']:.4e} cell-updates/s ({cpu['kind']}, 1 core); all cores: "
              f"{cpu['all_cores']['value']:.4e} on {cpu['all_cores']['cores']} threads "
              f"({cpu['all_cores']['cpu_model']}, identical {cpu['all_cores']['identical_to_serial']})",
              file=sys.stderr, flush=True)

    extras = {}
    if not a.no_extras:
        extras["batched_patches_fp32"] = run_batched("float")
        print(f"# batched: {extras['batched_patches_fp32']['value']:.4e} cell-updates/s", file=sys.stderr,
              flush=True)
        torch.cuda.empty_cache()
        extras["lax_friedrichs_512"] = run_lf(a.n)
        torch.cuda.empty_cache()
        extras["reinit_256"] = run_reinit()
        print(f"# reinit: fp64 {extras['reinit_256']['double']['ms_per_iteration']:.2f} ms/it, fp32 "
              f"{extras['reinit_256']['float']['ms_per_iteration']:.2f} ms/it", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        extras["config5_512_full"] = run_config5()
        c5 = extras["config5_512_full"]
        sp = c5["fp32_error_split"]
        print(f"# config5: fp32 {c5['float']['seconds']:.2f} s, fp64 {c5['double']['seconds']:.2f} s, near-interface "
              f"error {c5['fp32_error_vs_fp64']['max_near_interface_over_dx']:.3e} dx (inputs only "
              f"{sp['inputs_only']['max_over_dx']:.2e}, fp32 storage only {sp['storage_only']['max_over_dx']:.2e}, "
              f"phi alone {sp['storage_phi_only']['max_over_dx']:.2e}, s1 and s2 alone "
              f"{sp['storage_stages_only']['max_over_dx']:.2e}, one step {sp['one_step']['max_over_dx']:.2e})",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        extras["throughput_1024_fp32"] = run_throughput()
        print(f"# 1024^3: {extras['throughput_1024_fp32']['value']:.4e} cell-updates/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        extras["curvature_example"] = run_curvature(cpu=not a.no_cpu_baseline)
        torch.cuda.empty_cache()
        extras["curvature_local_example"] = run_curvature_local(cpu=not a.no_cpu_baseline)
        extras["field_extension_512"] = run_field_extension()
        torch.cuda.empty_cache()
        if not a.no_cpu_baseline:
            extras["config1_fp64_128"] = run_config1()
            c1 = extras["config1_fp64_128"]
            print(f"# config1: gpu {c1['gpu_s']:.2f} s, cpu {c1['cpu_s']:.1f} s, max rel diff "
                  f"{c1['max_rel_diff_vs_reference']:.2e}", file=sys.stderr, flush=True)
    adj = None
    if not a.no_adjoint:
        adj = run_adjoint(fd=not a.no_fd)
        print(f"# adjoint: fp32 {adj['adjoint_s']:.3f} s fwd+bwd, fp64 {adj['double']['forward_backward_s']:.3f} s, "
              f"fp32 grad err {adj['fp32_grad_rel_err_vs_fp64']}, fd {adj.get('fd_check_fp64', {}).get('max_rel_err')}",
              file=sys.stderr, flush=True)
        sp = adj["fp32_grad_split"]
        print("# fp32 gradient split (rel. err c / phi0): " + ", ".join(
            f"{k} {v['c']:.2e} / {v['phi0']:.2e}" for k, v in sp.items() if isinstance(v, dict)),
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    line = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"weno5_tvdrk3_advection_{a.n}cubed_periodic_64spheres_taylor_green",
                   "grid": [a.n, a.n, a.n], "ghost": 3, "cfl": 0.5, "dt": r32["dt"],
                   "tiling": r32["tiling"], "velocity": "stored Taylor-Green arrays",
                   "ghosts": a.ghosts, "fp32_upwind_zero_tol": SWITCH_TOL},
        "fp64": None if r64 is None else {"value": r64["value"], "ms_per_step": r64["ms_per_step"],
                                          "kernel_ms_avg": r64["kernel_ms_avg"],
                                          "achieved_gbs": r64["achieved_gbs"], "tiling": r64["tiling"]},
        "fp32_over_fp64": None if r64 is None else r32["value"] / r64["value"],
        "fp32_error_vs_fp64_one_step": err_one,
        "fp32_error_vs_fp64": err,
        "fp32_error_vs_fp64_switch_1e-5": err_build_tol,
        "fp32_error_split": split,
        "roofline": {"bound": "hbm", "achieved": r32["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r32["achieved_gbs"] / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": "k_stage (fused advection stage)",
                     "kernel_ms_avg": r32["kernel_ms_avg"],
                     "algorithmic_bytes_per_launch": r32["algorithmic_bytes_per_launch"],
                     "measured_peak_gbs": {k: v["gbs"] for k, v in probe.items() if "gbs" in v},
                     "frac_of_measured_read5_write1": r32["achieved_gbs"] / best5,
                     "frac_of_march_pattern": (r32["achieved_gbs"] / probe["march_pattern"]["gbs"]
                                               if "march_pattern" in probe else None),
                     "valu": valu_roofline(r32, valu_issue_peak(probe))},
        "cpu_baseline": cpu,
        "adjoint_s": None if adj is None else adj["adjoint_s"],
        "adjoint": adj,
        "extras": extras or None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
