"""GPU parity of every drop-in LSM3D_* entry point against the oracle.

Each routine is called through the C ABI of the HIP library and through the C
restatement with identical seeded arguments (uneven ghost boxes, non-zero
lower bounds, velocity zeros and sub-tolerance magnitudes, every boundary
face, both control-volume signs, both reinitialization sign options).

Tolerances (relative to the largest magnitude of the oracle's output):
  double: 1e-12 (the reference's FP64 parity bar); the kernels follow the
          Fortran operation order, so most routines agree to the last bits;
  float : 2e-5; the float build computes in FP32 throughout while the
          reference's float build promotes its double literals, so results
          differ by FP32 rounding; ENO stencil choices are identical because
          the undivided differences are the same FP32 subtractions.
The D1/D2/D3 scratch arrays are not compared: the HIP library does not write
them.
"""
import numpy as np
import pytest

from lsm_cases import cases_for
from lsmlib_amd import SIGNATURES, toolbox
from lsmlib_amd.signatures import BOX_ROUTINES
from oracle.oracle import oracle_lib

TOL = {"double": 1e-12, "float": 2e-5}
SUM_TOL = {"double": 1e-12, "float": 1e-5}


def is_sum(name):
    """Reductions that add over cells: the reference accumulates in
    LSMLIB_REAL cell by cell, the GPU in FP64 over blocks."""
    return "volume" in name or "surface" in name or "aveabsdiff" in name


def compare(name, real, case, got_args, got_ret, want_args, want_ret):
    tol = TOL[real]
    for (i, g), (_, w) in zip(got_args.arrays(), want_args.arrays()):
        if i in case.scratch:
            continue
        scale = max(float(np.abs(w).max()), 1e-30)
        err = float(np.abs(g.astype(np.float64) - w.astype(np.float64)).max()) / scale
        assert err <= tol, f"{name}[{real}] array arg {i}: rel err {err:.3e} > {tol}"
    if want_ret is not None:
        w = np.atleast_1d(np.asarray(want_ret, dtype=np.float64))
        g = np.atleast_1d(np.asarray(got_ret, dtype=np.float64))
        t = SUM_TOL[real] if is_sum(name) else tol
        err = float(np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-30)))
        assert err <= t, f"{name}[{real}] scalar: got {g} want {w} (rel {err:.3e})"


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("name", BOX_ROUTINES)
def test_host_memory_matches_oracle(name, real):
    lib, ora = toolbox(real), oracle_lib(real)
    for case in cases_for(name, real):
        a, b = case.copy(), case.copy()
        rw = ora.call(name, *a.args)
        rg = lib.call(name, *b.args)
        compare(name, real, case, b, rg, a, rw)


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("name", BOX_ROUTINES)
def test_device_memory_matches_host_memory(name, real, device):
    """Device pointers run in place and give the host-staged results bit for bit."""
    import torch
    lib = toolbox(real)
    case = cases_for(name, real)[0]
    host = case.copy()
    r_host = lib.call(name, *host.args)
    dev_args = [torch.from_numpy(a.copy()).to(device) if isinstance(a, np.ndarray) else a
                for a in case.args]
    r_dev = lib.call(name, *dev_args)
    torch.cuda.synchronize()
    for (i, h), d in zip(host.arrays(), [a for a in dev_args if hasattr(a, "data_ptr")]):
        if i in case.scratch:
            continue
        assert np.array_equal(h, d.cpu().numpy()), f"{name}[{real}] arg {i} differs"
    if r_host is not None:
        assert np.array_equal(np.atleast_1d(r_host), np.atleast_1d(r_dev))
