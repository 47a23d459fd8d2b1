"""The window construction of the full-size GPU parity tests
(tests/test_gpu_full_size.py), checked on the CPU: the oracle advancing a
42^3 window (a 24^3 box plus 9 cells per side, wrapped periodically, outer
layers read as stored ghosts) gives the inner 24^3 cells of the oracle's
whole-field step exactly -- for the periodic advection step (windows on the
interface and across all three periodic faces) and for the reinitialization
iteration (windows inside the fill box).  With a margin of 6 instead of 9
the windows would differ, which this also checks, so the margin is needed."""
import numpy as np
import torch

import tests.test_gpu_full_size as W
from lsmlib_amd.workloads import sphere_union_taylor_green, wavy_sphere_reinit
from oracle.oracle import oracle_lib
from oracle.steps import OracleStepper, fill_ghosts


def test_advection_windows_equal_whole_field_step():
    ora = oracle_lib("double")
    n = 64
    grid, phi, vel, bc = sphere_union_taylor_green(n, n_spheres=6, dtype=torch.float64)
    dt = 0.5 / n
    N = (n, n, n)
    origins = W.windows_near_interface(grid.interior(phi), grid.dx[0], 3, seed=7, wrap=True, n=N)
    origins.append((n - 11, n - 14, n - 10))
    ph_w = [W.window(grid.interior(phi), o, N) for o in origins]
    vel_w = [tuple(W.window(grid.interior(v), o, N) for v in vel) for o in origins]
    full = phi.numpy().copy()
    fill_ghosts(ora, full, grid, bc)
    OracleStepper(ora, grid, bc).advect_rk3_step(full, tuple(v.numpy().copy() for v in vel), dt)
    full_t = torch.from_numpy(full)
    op = OracleStepper(ora, W.window_grid(grid.dx), (0,) * 6)
    short = 0.0
    for o, ph, v in zip(origins, ph_w, vel_w):
        want_w = W.window(grid.interior(full_t), o, N)
        op.advect_rk3_step(ph, v, dt)
        assert np.array_equal(W.inner(ph), W.inner(want_w)), o
        m = W.MARGIN - 3  # the ring a 6-cell margin would have to get right
        short = max(short, float(np.abs(ph[m:-m, m:-m, m:-m] - want_w[m:-m, m:-m, m:-m]).max()))
    assert short > 0.0


def test_reinit_windows_equal_whole_field_iteration():
    ora = oracle_lib("double")
    grid, phi, bc = wavy_sphere_reinit(48, torch.float64, "cpu")
    N = (48, 48, 48)
    dt = 0.5 * grid.dx[0]
    origins = W.windows_near_interface(grid.interior(phi), grid.dx[0], 3, seed=3, wrap=False, n=N)
    ph_w = [W.window(grid.interior(phi), o, N) for o in origins]
    full = phi.numpy().copy()
    fill_ghosts(ora, full, grid, bc)
    OracleStepper(ora, grid, bc).reinit_rk3_iteration(full, dt, phi0=full.copy(), use_phi0=True)
    full_t = torch.from_numpy(full)
    op = OracleStepper(ora, W.window_grid(grid.dx), (0,) * 6)
    for o, ph in zip(origins, ph_w):
        op.reinit_rk3_iteration(ph, dt, phi0=ph.copy(), use_phi0=True)
        assert np.array_equal(W.inner(ph), W.inner(W.window(grid.interior(full_t), o, N))), o
