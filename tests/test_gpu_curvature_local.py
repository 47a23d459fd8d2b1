"""GPU parity of the localized curvature example (lsmlib_amd.curvature_local)
against its composition from the oracle's narrow-band routines
(oracle/steps.py LocalCurvatureFlowOracle), on the example's problem at
dx = 0.04 (38 x 38 x 25 cells, three-sphere mask).

FP64: the same field bit for bit, the same steps, reinitialization triggers
and interval maxima (the narrow-band kernels follow the Fortran's operation
order; the band lists are built in the reference's order).  The phase
volumes are sums (FP64 accumulation in parallel here, sequential in the
oracle): relative 1e-12.
FP32: against the float oracle over the first 25 steps, 1e-4 (the float
kernels compute in FP32 throughout; the reference's float build promotes its
double literals).
"""
import numpy as np
import pytest
import torch

from lsmlib_amd.workloads import pore_throat


@pytest.mark.gpu
@pytest.mark.parametrize("native", [True, False])
def test_local_curvature_run_fp64_bit_exact(device, native):
    """native: the step sequenced by lsm3d_amd_local_curvature_step; False:
    the same routines called one by one from Python."""
    from lsmlib_amd.curvature_local import LocalCurvatureFlow
    from oracle.oracle import oracle_lib
    from oracle.steps import LocalCurvatureFlowOracle
    g, phi, mask = pore_throat(0.04)
    pc = phi.numpy().copy()
    want = LocalCurvatureFlowOracle(oracle_lib("double"), g, mask.numpy().copy()).run(pc, max_outer=2)
    gg, pd, md = pore_throat(0.04, torch.float64, device)
    got = LocalCurvatureFlow(gg, "double", md, device=device).run(pd, max_outer=2, native=native)
    torch.cuda.synchronize()
    assert got["steps"] == want["steps"] and got["triggers"] == want["triggers"]
    assert got["reinit_iters"] == want["reinit_iters"]
    assert np.array_equal(pd.cpu().numpy(), pc)
    for (t1, e1, v1), (t2, e2, v2) in zip(got["history"], want["history"]):
        assert t1 == t2 and e1 == e2 and abs(v1 - v2) <= 1e-12 * abs(v2)


@pytest.mark.gpu
def test_local_curvature_steps_fp32(device):
    from lsmlib_amd.curvature_local import LocalCurvatureFlow
    from oracle.oracle import oracle_lib
    from oracle.steps import LocalCurvatureFlowOracle
    g, phi, mask = pore_throat(0.04, torch.float32)
    pc = phi.numpy().copy()
    LocalCurvatureFlowOracle(oracle_lib("float"), g, mask.numpy().copy()).run(pc, max_steps=25)
    gg, pd, md = pore_throat(0.04, torch.float32, device)
    LocalCurvatureFlow(gg, "float", md, device=device).run(pd, max_steps=25)
    torch.cuda.synchronize()
    assert float(np.abs(pd.cpu().numpy() - pc).max()) < 1e-4


@pytest.mark.gpu
def test_native_step_rejects_host_arrays(device):
    """The native step takes its arrays as device memory; a host array in the
    model is reported before any kernel runs."""
    import ctypes
    from lsmlib_amd._native import LsmError
    from lsmlib_amd.curvature_local import LocalCurvatureFlow
    gg, pd, md = pore_throat(0.04, torch.float64, device)
    flow = LocalCurvatureFlow(gg, "double", md, device=device)
    m = flow.model()
    host = np.zeros(flow.grid.shape)
    m.phi0 = host.ctypes.data
    with pytest.raises(LsmError, match="device memory"):
        flow.lib.fused("lsm3d_amd_local_curvature_step", ctypes.byref(m), pd.data_ptr(), 0.0, 0.1)
    m = flow.model()   # a valid model afterwards steps normally
    dt = flow.lib.fused("lsm3d_amd_local_curvature_step", ctypes.byref(m), pd.data_ptr(), 0.0, 0.1)
    assert 0 < dt <= 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("masked", [True, False])
def test_fused_step_equals_routine_chain(device, real, masked):
    """lsm3d_amd_local_curvature_step (the two stages fused per level-0 point,
    dt chosen on the device) against the same step sequenced routine by routine
    (*LOCAL kernels, dt on the host), at dx = 0.02 (75 x 75 x 50 cells), with
    and without the mask, 6 steps from one initial state: phi, the stage
    fields, dt and the level limits equal bit for bit in both precisions."""
    from lsmlib_amd.curvature_local import LocalCurvatureFlow
    dtype = torch.float64 if real == "double" else torch.float32
    gg, phi, md = pore_throat(0.02, dtype, device)
    runs = []
    for native in (True, False):
        f = LocalCurvatureFlow(gg, real, md if masked else None, device=device)
        p = phi.clone()
        f.box.fill_ghosts(p, masked=False)
        f.reinitialize_box(p, f.gamma + gg.dx[0])
        dts, sub = [], 0.0
        for _ in range(6):
            dt = f.step_native(p, sub, 0.1) if native else f.step(p, sub, 0.1)
            dts.append(dt)
            sub += dt
        torch.cuda.synchronize()
        runs.append((p, f.s1.clone(), f.nxt.clone(), dts, f.n_lo.copy(), f.n_hi.copy()))
    (p1, a1, b1, d1, lo1, hi1), (p2, a2, b2, d2, lo2, hi2) = runs
    assert d1 == d2
    assert np.array_equal(lo1, lo2) and np.array_equal(hi1, hi2)
    assert torch.equal(p1, p2) and torch.equal(a1, a2) and torch.equal(b1, b2)
