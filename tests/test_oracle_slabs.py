"""The slab-parallel CPU baseline stepper (oracle.steps.SlabOracleStepper)
against the serial composition (OracleStepper): the same field bit for bit
after three WENO5 / TVD-RK3 steps, with the C restatement and, where it is
built, the flang build of the reference Fortran (whose routines must then be
safe to call from several threads at once)."""
import numpy as np
import pytest
import torch

from lsmlib_amd.workloads import sphere_union_taylor_green
from oracle.oracle import oracle_lib, ref_lib
from oracle.steps import OracleStepper, SlabOracleStepper, fill_ghosts


@pytest.mark.parametrize("which", ["port", "reference"])
@pytest.mark.parametrize("real", ["float", "double"])
def test_slab_parallel_equals_serial(which, real):
    lib = oracle_lib(real) if which == "port" else ref_lib(real)
    if lib is None:
        pytest.skip("reference build not present")
    dtype = torch.float32 if real == "float" else torch.float64
    grid, phi, vel, bc = sphere_union_taylor_green(28, dtype=dtype, device="cpu")
    vn = tuple(v.numpy().copy() for v in vel)
    a = phi.numpy().copy()
    fill_ghosts(lib, a, grid, bc)
    b = a.copy()
    dt = lib.lsm3dcomputestableadvectiondt_(vn[0], vn[1], vn[2], grid.gb, grid.fb, *grid.dx, 0.5)
    ser, par = OracleStepper(lib, grid, bc), SlabOracleStepper(lib, grid, bc, threads=5)
    try:
        for _ in range(3):
            ser.advect_rk3_step(a, vn, dt)
            par.advect_rk3_step(b, vn, dt)
    finally:
        par.close()
    assert np.array_equal(a, b)
