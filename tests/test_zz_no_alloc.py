"""The allocation-free CFL reduction against LSM3D_COMPUTE_STABLE_ADVECTION_DT.

lsm3d_amd_compute_stable_advection_dt_noalloc keeps its partial maxima in a
fixed device array instead of the library's grown scratch.  A maximum does
not depend on the order of the partials, so its dt must equal the drop-in
routine's bit for bit: checked on two calls in a row, the second over a
smaller box (stale partials of the first call must not leak in), in both
precisions, plus the empty-box value and the size limit.
"""
import ctypes

import numpy as np
import pytest
import torch

from lsmlib_amd import toolbox
from lsmlib_amd._native import LsmError
from lsmlib_amd.pylsmlib import box_c

pytestmark = pytest.mark.gpu

DT = {"double": torch.float64, "float": torch.float32}


def _vel(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    return tuple(torch.from_numpy(rng.uniform(-2.0, 2.0, size=shape)).to(dtype=dtype, device="cuda").contiguous()
                 for _ in range(3))


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(lib, real, vel, gb, ib, h, cfl):
    out = torch.full((1,), float("nan"), dtype=DT[real], device="cuda")
    lib.fused("lsm3d_amd_compute_stable_advection_dt_noalloc", _p(out), *map(_p, vel), box_c(gb), box_c(ib), *h, cfl)
    torch.cuda.synchronize()
    return float(out.item())


@pytest.mark.parametrize("real", ["double", "float"])
def test_noalloc_dt_equals_the_drop_in_routine_on_two_calls(real):
    lib = toolbox(real)
    h = (1.0 / 70, 1.0 / 45, 1.0 / 33)
    # call 1: 70 x 45 x 33 include box on a shifted ghost box
    gb1 = (-3, 72, -2, 46, -4, 35)
    ib1 = (0, 69, 0, 44, 0, 32)
    shape1 = (gb1[5] - gb1[4] + 1, gb1[3] - gb1[2] + 1, gb1[1] - gb1[0] + 1)
    v1 = _vel(shape1, 1, DT[real])
    want1 = lib.lsm3dcomputestableadvectiondt_(*v1, gb1, ib1, *h, 0.5)
    got1 = _noalloc(lib, real, v1, gb1, ib1, h, 0.5)
    assert got1 == want1
    # call 2: a smaller box and other values (partials of call 1 stay in the array)
    gb2 = (0, 20, 0, 16, 0, 8)
    ib2 = (2, 18, 1, 15, 0, 8)
    shape2 = (9, 17, 21)
    v2 = _vel(shape2, 2, DT[real])
    want2 = lib.lsm3dcomputestableadvectiondt_(*v2, gb2, ib2, *h, 0.8)
    got2 = _noalloc(lib, real, v2, gb2, ib2, h, 0.8)
    assert got2 == want2
    assert want1 != want2


@pytest.mark.parametrize("real", ["double", "float"])
def test_noalloc_dt_of_an_empty_box_is_the_reference_value(real):
    lib = toolbox(real)
    v = _vel((4, 4, 4), 3, DT[real])
    gb = (0, 3, 0, 3, 0, 3)
    ib = (2, 1, 0, 3, 0, 3)  # ihi < ilo
    want = lib.lsm3dcomputestableadvectiondt_(*v, gb, ib, 0.1, 0.1, 0.1, 0.5)
    assert _noalloc(lib, real, v, gb, ib, (0.1, 0.1, 0.1), 0.5) == want


def test_noalloc_dt_rejects_a_box_beyond_the_fixed_scratch():
    lib = toolbox("double")
    v = _vel((1, 1, 64), 4, torch.float64)
    big = (0, 64 * 65537 - 1, 0, 3, 0, 0)  # ceil(nx/64) * ceil(ny/4) = 65537 partials per plane
    with pytest.raises(LsmError, match="too large"):
        _noalloc(lib, "double", v, big, big, (1.0, 1.0, 1.0), 0.5)
