"""GPU parity of the fused device stages against the oracle's toolbox chain.

The oracle advances the same seeded fields with the reference's routine
sequence (oracle/steps.py); the GPU advances them with one fused kernel
per Runge-Kutta stage.  FP64 must agree to 1e-12 relative to max|phi| (the
reference's FP64 parity bar).  FP32 is compared with the FP64 oracle: its
error is reported and bounded by 1e-5 * max|phi| globally and 1e-5 * dx near
the interface (|phi| < 3 dx) over the few steps run here.
"""
import ctypes

import numpy as np
import pytest
import torch

from lsmlib_amd import toolbox
from lsmlib_amd.pylsmlib import (BC_PERIODIC, BC_SIGNED_LINEAR, BC_LINEAR, BC_NEUMANN, Grid, Stepper,
                                 stable_advection_dt)
from lsmlib_amd.workloads import rotating_sphere, sphere_union, taylor_green, wavy_sphere_reinit
from oracle.oracle import oracle_lib
from oracle.steps import OracleStepper, fill_ghosts

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _np(t):
    return t.detach().cpu().numpy().copy()


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() / np.abs(b).max())


def small_periodic(n=(24, 20, 18), dtype=torch.float64):
    grid = Grid(n, (0.0, 0.0, 0.0), (1.0, 0.9, 0.8), 3)
    phi = sphere_union(grid, n_spheres=5, seed=11, rmin=0.15, rmax=0.3, dtype=dtype)
    vel = taylor_green(grid, dtype)
    return grid, phi, vel


def run_both(grid, phi0, vel, bc, steps, real, dt, mode="advect", use_phi0=False, vel_const=None,
             ghosts="synth", zero_tol=None):
    dev = torch.device("cuda:0")
    st = Stepper(grid, real, bc, ghosts=ghosts, zero_tol=zero_tol)
    phi = phi0.to(dev).clone()
    s1, s2 = st.empty(), st.empty()
    v = vel_const if vel_const is not None else (tuple(a.to(dev) for a in vel) if vel is not None else None)
    p0 = phi.clone() if mode == "reinit" else None
    st.fill_ghosts(phi)
    for _ in range(steps):
        if mode == "advect":
            st.advect_rk3_step(phi, s1, s2, v, dt)
        elif mode == "advect_rk2":
            st.advect_rk2_step(phi, s1, v, dt)
        else:
            st.reinit_rk3_iteration(phi, s1, s2, dt, phi0=p0, use_phi0=use_phi0)
    torch.cuda.synchronize()

    ora = oracle_lib("double")
    op = OracleStepper(ora, grid, bc)
    ph = phi0.to(torch.float64).numpy().copy()
    if vel_const is not None:
        vn = tuple(np.full(grid.shape, c, dtype=np.float64) for c in vel_const)
    elif vel is None:
        vn = None
    else:
        vn = tuple(a.to(torch.float64).numpy().copy() for a in vel)
    q0 = ph.copy()
    fill_ghosts(ora, ph, grid, bc)
    for _ in range(steps):
        if mode == "advect":
            op.advect_rk3_step(ph, vn, dt)
        elif mode == "advect_rk2":
            op.advect_rk2_step(ph, vn, dt)
        else:
            op.reinit_rk3_iteration(ph, dt, phi0=q0, use_phi0=use_phi0)
    return grid.interior(_np(phi)), grid.interior(ph)


@pytest.mark.parametrize("ghosts", ["synth", "fill", "images"])
def test_advect_rk3_periodic_fp64(ghosts):
    grid, phi, vel = small_periodic()
    dt = 0.5 * min(grid.dx)
    got, want = run_both(grid, phi, vel, (BC_PERIODIC,) * 6, 4, "double", dt, ghosts=ghosts)
    assert _rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("ghosts", ["synth", "fill", "images"])
def test_advect_rk3_signed_linear_fp64(ghosts):
    grid, phi, vel, bc = rotating_sphere(20)
    dt = 0.5 * grid.dx[0] / 1.8
    got, want = run_both(grid, phi, vel, bc, 4, "double", dt, ghosts=ghosts)
    assert _rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("real", ["double", "float"])
def test_synthesised_ghosts_equal_filled_ghosts(real):
    """Stages that synthesise ghost values give bit-identical steps to stages
    that read ghosts filled by the device extrapolation kernel."""
    grid, phi, vel, bc = rotating_sphere(19)
    dt = 0.01
    dt_ = torch.float64 if real == "double" else torch.float32
    outs = []
    for ghosts in ("synth", "fill", "images"):
        outs.append(run_both(grid, phi.to(dt_), tuple(v.to(dt_) for v in vel), bc, 2, real, dt, ghosts=ghosts)[0])
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("ghosts", ["synth", "images"])
@pytest.mark.parametrize("bc", [BC_LINEAR, BC_NEUMANN])
def test_advect_rk3_other_bcs_fp64(bc, ghosts):
    grid, phi, vel, _ = rotating_sphere(16)
    dt = 0.4 * grid.dx[0]
    got, want = run_both(grid, phi, vel, (bc,) * 6, 3, "double", dt, ghosts=ghosts)
    assert _rel(got, want) <= FP64_TOL


def test_advect_rk2_fp64():
    grid, phi, vel = small_periodic()
    got, want = run_both(grid, phi, vel, (BC_PERIODIC,) * 6, 3, "double", 0.4 * min(grid.dx), mode="advect_rk2")
    assert _rel(got, want) <= FP64_TOL


def test_advect_const_velocity_fp64():
    grid, phi, vel = small_periodic()
    got, want = run_both(grid, phi, vel, (BC_PERIODIC,) * 6, 3, "double", 0.3 * min(grid.dx),
                         vel_const=(0.7, -0.4, 0.25))
    assert _rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("ghosts", ["synth", "images"])
@pytest.mark.parametrize("use_phi0", [False, True])
def test_reinit_rk3_fp64(use_phi0, ghosts):
    grid, phi, bc = wavy_sphere_reinit(20)
    got, want = run_both(grid, phi, None, bc, 3, "double", 0.5 * grid.dx[0], mode="reinit",
                         use_phi0=use_phi0, ghosts=ghosts)
    assert _rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("real", ["double", "float"])
def test_face_ghost_writes_for_every_tiling(real):
    """Stages that write their output's extrapolated face ghosts (and read
    stored ghosts) step the fill box exactly as stages that synthesise the
    ghost values, for every tile shape and z-chunk -- including the tilings
    whose last tile or chunk is one cell wide, where the face ghosts come from
    a ghost fill after the stage -- and the face ghost layers they write equal
    those of the reference's fill."""
    dev = torch.device("cuda:0")
    dt_ = torch.float64 if real == "double" else torch.float32
    grid = Grid((37, 29, 23), (0.0, 0.0, 0.0), (1.0, 0.8, 0.6), 3)
    phi = sphere_union(grid, n_spheres=4, seed=7, rmin=0.12, rmax=0.25, dtype=dt_).to(dev)
    vel = tuple(v.to(dev) for v in taylor_green(grid, dt_))
    bc = (BC_SIGNED_LINEAR, BC_LINEAR, BC_SIGNED_LINEAR, BC_NEUMANN, BC_LINEAR, BC_SIGNED_LINEAR)
    lib = toolbox(real)
    ref_st = Stepper(grid, real, bc, ghosts="synth", tiling=(2, 8))
    p = phi.clone()
    s1, s2 = ref_st.empty(), ref_st.empty()
    ref_st.advect_rk3_step(p, s1, s2, vel, 0.01)
    ref_st.reinit_rk3_iteration(p, s1, s2, 0.005)
    ref = grid.interior(_np(p))
    st = Stepper(grid, real, bc, ghosts="images", tiling=None)  # the loop below sets the tiling
    try:
        for cfg in range(lib.lib.lsm3d_amd_stage_tiling_count()):
            for kz in (1, 7, 64):
                lib.lib.lsm3d_amd_set_stage_tiling(cfg, kz)
                p = phi.clone()
                st.fill_ghosts(p)
                s1, s2 = st.empty(), st.empty()
                st.advect_rk3_step(p, s1, s2, vel, 0.01)
                st.reinit_rk3_iteration(p, s1, s2, 0.005)
                torch.cuda.synchronize()
                assert np.array_equal(grid.interior(_np(p)), ref), f"tiling {cfg}/{kz} differs"
                # the face ghost layers written by the last stage equal a fresh fill's
                q = p.clone()
                st.fill_ghosts(q)
                torch.cuda.synchronize()
                a, b = _np(p), _np(q)
                g = 3
                gi, gj, gk = (g + grid.n[0], g + grid.n[1], g + grid.n[2])
                assert np.array_equal(a[g:gk, g:gj, :], b[g:gk, g:gj, :]), f"tiling {cfg}/{kz}: x face ghosts"
                assert np.array_equal(a[g:gk, :, g:gi], b[g:gk, :, g:gi]), f"tiling {cfg}/{kz}: y face ghosts"
                assert np.array_equal(a[:, g:gj, g:gi], b[:, g:gj, g:gi]), f"tiling {cfg}/{kz}: z face ghosts"
    finally:
        lib.lib.lsm3d_amd_set_stage_tiling(0, 32)


@pytest.mark.parametrize("ghosts", ["synth", "images"])
def test_advect_fp32_vs_fp64_oracle(ghosts):
    grid, phi, vel = small_periodic()
    dt = 0.5 * min(grid.dx)
    got, want = run_both(grid, phi.float(), tuple(v.float() for v in vel), (BC_PERIODIC,) * 6, 4, "float", dt,
                         ghosts=ghosts)
    err = np.abs(got.astype(np.float64) - want)
    near = np.abs(want) < 3 * min(grid.dx)
    assert err.max() <= 1e-5 * np.abs(want).max()
    assert err[near].max() <= 1e-5 * min(grid.dx)


@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("ghosts", ["images", "synth"])
def test_padded_rows_step_identically(real, ghosts):
    """A grid whose x ghost region is padded so rows start on 128-byte lines
    (Grid(align=...)) steps bit-identically to the plain 3-cell ghost layout."""
    dt_ = torch.float64 if real == "double" else torch.float32
    outs = []
    for align in (0, 32):
        grid = Grid((40, 36, 24), (0.0, 0.0, 0.0), (1.0, 0.9, 0.6), 3, align)
        phi = sphere_union(grid, n_spheres=5, seed=11, rmin=0.15, rmax=0.3, dtype=dt_).cuda()
        vel = tuple(v.cuda() for v in taylor_green(grid, dt_))
        st = Stepper(grid, real, (BC_PERIODIC,) * 6, ghosts=ghosts)
        st.fill_ghosts(phi)
        s1, s2 = st.empty(), st.empty()
        for _ in range(3):
            st.advect_rk3_step(phi, s1, s2, vel, 0.4 * min(grid.dx))
        torch.cuda.synchronize()
        outs.append(grid.interior(phi).cpu())
    assert torch.equal(outs[0], outs[1])


def test_fp32_upwind_switch_threshold():
    """Velocities below the float build's LSMLIB_ZERO_TOL (1e-5) but above the
    double build's (1e-11): by default the FP32 stages add no advection term
    there (the float build's behaviour); with zero_tol=1e-11 they follow the
    FP64 oracle to FP32 rounding."""
    grid, phi, vel = small_periodic()
    slow = tuple((9e-6 * v).float() for v in vel)  # |V| < 1e-5 everywhere
    dt, steps = 5.0, 3
    phi32 = phi.float()
    got_default, want = run_both(grid, phi32, slow, (BC_PERIODIC,) * 6, steps, "float", dt)
    start = grid.interior(phi32.numpy()).astype(np.float64)
    moved = np.abs(want - start).max()
    scale = np.abs(want).max()
    assert moved > 1e-4
    assert np.abs(got_default.astype(np.float64) - start).max() <= 1e-6 * scale
    got, want2 = run_both(grid, phi32, slow, (BC_PERIODIC,) * 6, steps, "float", dt, zero_tol=1e-11)
    assert np.array_equal(want, want2)
    assert np.abs(got.astype(np.float64) - want).max() <= 1e-6 * scale
    lib = Stepper(grid, "float", (BC_PERIODIC,) * 6).lib
    assert lib.zero_tol() == pytest.approx(1e-11)  # the setting is per host thread and persists
    lib.set_zero_tol(None)
    assert lib.zero_tol() == pytest.approx(1e-5)


def test_reinit_fp32_vs_fp64_oracle():
    grid, phi, bc = wavy_sphere_reinit(20)
    got, want = run_both(grid, phi.float(), None, bc, 3, "float", 0.5 * grid.dx[0], mode="reinit",
                         use_phi0=True)
    err = np.abs(got.astype(np.float64) - want)
    assert err.max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("real", ["double", "float"])
def test_every_tiling_gives_identical_results(real):
    """Tile shape and z-chunk change only the work decomposition."""
    dev = torch.device("cuda:0")
    grid, phi, vel = small_periodic((37, 29, 23), torch.float64 if real == "double" else torch.float32)
    lib = toolbox(real)
    st = Stepper(grid, real, ghosts="synth", tiling=None)  # the loop below sets the tiling
    vel = tuple(v.to(dev) for v in vel)
    ref = None
    try:
        for cfg in range(lib.lib.lsm3d_amd_stage_tiling_count()):
            for kz in (1, 7, 64):
                lib.lib.lsm3d_amd_set_stage_tiling(cfg, kz)
                p = phi.to(dev).clone()
                s1, s2 = st.empty(), st.empty()
                st.advect_rk3_step(p, s1, s2, vel, 0.01)
                st.reinit_rk3_iteration(p, s1, s2, 0.005)
                st.advect_lf_rk3_step(p, s1, s2, vel, 0.01)
                torch.cuda.synchronize()
                out = _np(p)
                assert np.isfinite(out).all(), f"tiling {cfg}/{kz} left non-finite values"
                if ref is None:
                    ref = out
                else:
                    assert np.array_equal(out, ref), f"tiling {cfg}/{kz} differs"
    finally:
        lib.lib.lsm3d_amd_set_stage_tiling(0, 32)


def test_dropin_chain_on_device_matches_fused():
    """The plain drop-in path (toolbox entry points on device pointers, the
    reference's call sequence) and the fused stage give the same step."""
    dev = torch.device("cuda:0")
    grid, phi, vel, bc = rotating_sphere(18)
    dx, dy, dz = grid.dx
    dt = 0.01
    lib = toolbox("double")
    gb, fb = grid.gb, grid.fb
    v = tuple(a.to(dev) for a in vel)
    p = phi.to(dev).clone()
    z = lambda: torch.zeros(grid.shape, dtype=torch.float64, device=dev)  # noqa: E731
    gx, gy, gz, rhs, D1, s1, s2, new = (z() for _ in range(8))

    def L(ps):
        lib.lsm3dupwindhjweno5_(gx, gy, gz, gb, ps, gb, v[0], v[1], v[2], gb, D1, gb, fb, dx, dy, dz)
        lib.lsm3dzerooutlevelseteqnrhs_(rhs, gb)
        lib.lsm3daddadvectiontermtolserhs_(rhs, gb, gx, gy, gz, gb, v[0], v[1], v[2], gb, fb)
        return rhs

    def fill(a):
        for face in range(6):
            lib.lsm3dsignedlinearextrapolation_(a, gb, fb, face)

    fill(p)
    lib.lsm3dtvdrk3stage1_(s1, gb, p, gb, L(p), gb, fb, dt)
    fill(s1)
    lib.lsm3dtvdrk3stage2_(s2, gb, s1, gb, p, gb, L(s1), gb, fb, dt)
    fill(s2)
    lib.lsm3dtvdrk3stage3_(new, gb, s2, gb, p, gb, L(s2), gb, fb, dt)

    st = Stepper(grid, "double", bc)
    q = phi.to(dev).clone()
    st.fill_ghosts(q)
    st.advect_rk3_step(q, st.empty(), st.empty(), v, dt)
    torch.cuda.synchronize()
    a, b = grid.interior(_np(new)), grid.interior(_np(q))
    assert _rel(a, b) <= FP64_TOL


def test_stable_dt_and_mass_on_device():
    dev = torch.device("cuda:0")
    grid, phi, vel, _ = rotating_sphere(24)
    v = tuple(a.to(dev) for a in vel)
    dt = stable_advection_dt(v, grid, 0.5)
    ora = oracle_lib("double")
    want = ora.lsm3dcomputestableadvectiondt_(*(a.numpy() for a in vel), grid.gb, grid.fb, *grid.dx, 0.5)
    assert dt == pytest.approx(want, rel=1e-15)
