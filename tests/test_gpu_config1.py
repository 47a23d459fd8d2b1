"""Config 1 at its full size on the GPU: 128^3 on [-1,1]^3, the sphere of
radius 0.35 at (0.35, 0, 0) in the rigid rotation u = (-y, x, 0),
signed-linear ghosts, HJ-WENO5 + TVD-RK3 at CFL 0.5, 100 steps in FP64,
against the oracle's composition of the reference chain
(LSM3D_UPWIND_HJ_WENO5 -> ZERO_OUT -> ADD_ADVECTION_TERM -> TVD_RK3_STAGE*,
with the signed-linear extrapolation between stages) over the same steps:
the flang build of the reference where it is present, else the C
restatement, called on k-slabs from several host threads (bit-identical to
the serial composition, tests/test_oracle_slabs.py).  The fused FP64 stages
keep the reference's operation order, so the field after 100 steps is
asserted bit for bit (the bar of 1e-12 relative to max|phi| would allow
more)."""
import os

import numpy as np
import pytest
import torch

from lsmlib_amd.pylsmlib import Stepper, stable_advection_dt
from lsmlib_amd.workloads import rotating_sphere
from oracle.oracle import oracle_lib, ref_lib
from oracle.steps import SlabOracleStepper, fill_ghosts

pytestmark = pytest.mark.gpu


def test_config1_128_fp64_100_steps_bit_identical_to_the_oracle(device):
    n, steps = 128, 100
    grid, phi, vel, bc = rotating_sphere(n, torch.float64, device)
    st = Stepper(grid, "double", bc)
    dt = stable_advection_dt(vel, grid, 0.5)
    p = phi.clone()
    s1, s2 = st.empty(), st.empty()
    for _ in range(steps):
        st.advect_rk3_step(p, s1, s2, vel, dt)
    torch.cuda.synchronize()
    got = grid.interior(p).cpu().numpy()
    lib = ref_lib("double") or oracle_lib("double")
    ph = phi.cpu().numpy().copy()
    vn = tuple(v.cpu().numpy().copy() for v in vel)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    op = SlabOracleStepper(lib, grid, bc, threads)
    fill_ghosts(lib, ph, grid, bc)
    try:
        for _ in range(steps):
            op.advect_rk3_step(ph, vn, dt)
    finally:
        op.close()
    want = grid.interior(ph)
    assert float(np.abs(got - want).max()) <= 1e-12 * float(np.abs(want).max())
    assert np.array_equal(got, want)
