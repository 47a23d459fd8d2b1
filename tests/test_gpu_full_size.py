"""Parity at the benchmarked sizes, on sampled windows.

One TVD-RK3 step moves information at most 9 cells (three stages of the
7-point WENO5 stencil), so the step's result inside a box depends only on
the input within 9 cells of it.  The GPU advances the whole field (the
512^3 sphere-union / Taylor-Green advection step with stored periodic ghost
images and the tuned tiling; the 256^3 reinitialization iteration with
stored face ghosts); the oracle (the reference's routine sequence,
oracle/steps.py) advances 42^3 windows of the same input -- a 24^3 box plus
9 cells per side, periodic wrap included -- with the window's outer 3
layers as stored ghosts (bc code 0: nothing is filled), and the inner 24^3
cells of its result are exact.  FP64: within 1e-12 of max|phi| of the
window; FP32 against the FP64 oracle: 1e-5 dx near the interface
(|phi| < 3 dx), 1e-5 max|phi| elsewhere.
"""
import numpy as np
import pytest
import torch

from lsmlib_amd.pylsmlib import Grid, Stepper
from lsmlib_amd.workloads import sphere_union_taylor_green, wavy_sphere_reinit
from oracle.oracle import oracle_lib
from oracle.steps import OracleStepper

pytestmark = pytest.mark.gpu

BOX, MARGIN = 24, 9
WIN = BOX + 2 * MARGIN


def windows_near_interface(phi_int, dx, count, seed, wrap, n):
    """Window origins (corner of the 42^3 window in interior indices, k, j, i):
    `count` centred on cells with |phi| < dx, drawn with `seed`; without wrap
    the windows stay inside the fill box."""
    near = torch.nonzero(phi_int.abs() < dx)
    g = torch.Generator().manual_seed(seed)
    pick = near[torch.randint(len(near), (count,), generator=g)].cpu().numpy()
    out = []
    for c in pick:
        o = c - WIN // 2
        if not wrap:
            o = np.clip(o, 0, np.array(n) - WIN)
        out.append(tuple(int(v) for v in o))
    return out


def window(t_int, o, n):
    """t_int[o + 0..41] along each axis, wrapping periodically, as FP64 numpy."""
    idx = [torch.arange(o[d], o[d] + WIN, device=t_int.device) % n[d] for d in range(3)]
    w = t_int.index_select(0, idx[0]).index_select(1, idx[1]).index_select(2, idx[2])
    return w.to(torch.float64).cpu().numpy().copy()


def inner(a):
    return a[MARGIN:MARGIN + BOX, MARGIN:MARGIN + BOX, MARGIN:MARGIN + BOX]


def window_grid(dx):
    nf = WIN - 6
    return Grid((nf, nf, nf), (0.0, 0.0, 0.0), tuple(nf * h for h in dx), 3, spacing=tuple(dx))


def check(got, want, dx, real, where):
    if real == "double":
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), where
    else:
        near = np.abs(want) < 3 * dx
        err = np.abs(got - want)
        assert err[near].max(initial=0.0) <= 1e-5 * dx, where
        assert err.max() <= 1e-5 * np.abs(want).max(), where


@pytest.mark.parametrize("real", ["double", "float"])
def test_headline_step_512_on_windows(real):
    dev = torch.device("cuda:0")
    dtype = torch.float64 if real == "double" else torch.float32
    grid, phi64, vel64, bc = sphere_union_taylor_green(512, dtype=torch.float64, device=dev,
                                                       align=16 if real == "double" else 32)
    dt = 0.0009765625  # CFL 0.5 with max |V| = 1 at dx = 1/512
    n = (512, 512, 512)
    origins = windows_near_interface(grid.interior(phi64), grid.dx[0], 3, seed=7, wrap=True, n=n)
    origins.append((512 - MARGIN - 2, 512 - MARGIN - 5, 512 - MARGIN - 1))  # across all three periodic faces
    # oracle inputs: windows of the FP64 fields (the FP32 run starts from these rounded once)
    ph_w = [window(grid.interior(phi64), o, n) for o in origins]
    vel_w = [tuple(window(grid.interior(v), o, n) for v in vel64) for o in origins]
    phi = phi64.to(dtype).contiguous()
    vel = tuple(v.to(dtype).contiguous() for v in vel64)
    del phi64, vel64
    st = Stepper(grid, real, bc, ghosts="images", zero_tol=1e-11)
    st.fill_ghosts(phi)
    s1, s2 = st.empty(), st.empty()
    print(f"512^3 {real} step, windows {origins}", flush=True)
    st.advect_rk3_step(phi, s1, s2, vel, dt)
    torch.cuda.synchronize()
    got_w = [inner(window(grid.interior(phi), o, n)) for o in origins]
    del phi, vel, s1, s2
    torch.cuda.empty_cache()
    wg = window_grid(grid.dx)
    ora = oracle_lib("double")
    op = OracleStepper(ora, wg, (0,) * 6)
    for o, ph, v, got in zip(origins, ph_w, vel_w, got_w):
        op.advect_rk3_step(ph, v, dt)
        check(got, inner(ph), grid.dx[0], real, f"window at {o}")


@pytest.mark.parametrize("real", ["double", "float"])
def test_reinit_iteration_256_on_windows(real):
    dev = torch.device("cuda:0")
    dtype = torch.float64 if real == "double" else torch.float32
    grid, phi64, bc = wavy_sphere_reinit(256, torch.float64, dev, align=16 if real == "double" else 32)
    n = (256, 256, 256)
    dt = 0.5 * grid.dx[0]
    origins = windows_near_interface(grid.interior(phi64), grid.dx[0], 4, seed=3, wrap=False, n=n)
    ph_w = [window(grid.interior(phi64), o, n) for o in origins]
    phi = phi64.to(dtype).contiguous()
    del phi64
    st = Stepper(grid, real, bc, ghosts="images", zero_tol=1e-11)  # the FP64 oracle's cut-off
    st.fill_ghosts(phi)
    phi0 = phi.clone()
    s1, s2 = st.empty(), st.empty()
    print(f"256^3 {real} reinitialization, windows {origins}", flush=True)
    st.reinit_rk3_iteration(phi, s1, s2, dt, phi0=phi0, use_phi0=True)
    torch.cuda.synchronize()
    got_w = [inner(window(grid.interior(phi), o, n)) for o in origins]
    wg = window_grid(grid.dx)
    ora = oracle_lib("double")
    op = OracleStepper(ora, wg, (0,) * 6)
    for o, ph, got in zip(origins, ph_w, got_w):
        p0 = ph.copy()
        op.reinit_rk3_iteration(ph, dt, phi0=p0, use_phi0=True)
        check(got, inner(ph), grid.dx[0], real, f"window at {o}")


def test_lf_step_512_equals_upwind_step_fp64():
    """Local Lax-Friedrichs (alpha_a = |v_a|) equals the upwind term wherever
    |v_a| >= the switch threshold (1e-11 in FP64; the Taylor-Green field has
    no cell below it), so two WENO5 / TVD-RK3 steps of the fused LF stage and
    of the fused upwind stage give the same 512^3 field bit for bit."""
    dev = torch.device("cuda:0")
    grid, phi, vel, bc = sphere_union_taylor_green(512, dtype=torch.float64, device=dev, align=16)
    dt = 0.0009765625
    st = Stepper(grid, "double", bc, ghosts="images")
    st.fill_ghosts(phi)
    up = phi.clone()
    s1, s2 = st.empty(), st.empty()
    for _ in range(2):
        st.advect_lf_rk3_step(phi, s1, s2, vel, dt)
        st.advect_rk3_step(up, s1, s2, vel, dt)
    torch.cuda.synchronize()
    assert torch.equal(grid.interior(phi), grid.interior(up))
