"""GPU parity of the curvature-model path against the oracle.

The fused kernels (include/lsm3d_amd.h, "curvature model") are compared with
the oracle's toolbox composition of the curvature example
(oracle/steps.py:CurvatureFlowOracle) on the example's own problem (the
pore throat between three spheres, with its mask) and on seeded noisy
fields with uneven ghost boxes.

Tolerances: FP64 kernels round every operation as the reference does
(contraction off), so stages, ghost fills and the stable dt agree with the
oracle bit for bit; whole runs are held to 1e-12 relative to max|phi|.  FP32
stages are compared with the float oracle to 2e-6 of max|phi|, whole runs
to 1e-4.
"""
import numpy as np
import pytest
import torch

from lsmlib_amd import toolbox
from lsmlib_amd.curvature import CurvatureFlow
from lsmlib_amd.pylsmlib import RK1, TVDRK2_STAGE2, Grid
from lsmlib_amd.workloads import pore_throat
from oracle.oracle import oracle_lib
from oracle.steps import CurvatureFlowOracle, impose_mask

pytestmark = pytest.mark.gpu

DT = {"double": torch.float64, "float": torch.float32}
NP = {"double": np.float64, "float": np.float32}


def _np(t):
    return t.detach().cpu().numpy().copy()


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() / max(np.abs(b).max(), 1e-300))


def noisy_case(real, n=(29, 21, 17), seed=3):
    """A smooth level set with noise on an uneven grid, and a mask that is
    active on part of the domain (including ghost cells)."""
    rng = np.random.default_rng(seed)
    grid = Grid(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.65), 3)
    X, Y, Z = (c.numpy() for c in grid.centers())
    phi = np.sqrt((X - 0.45) ** 2 + (Y - 0.4) ** 2 + (Z - 0.3) ** 2) - 0.25 + 0.004 * rng.standard_normal(X.shape)
    mask = 0.12 - np.sqrt((X - 0.9) ** 2 + (Y - 0.1) ** 2 + (Z - 0.6) ** 2) * 0.8
    return grid, phi.astype(NP[real]), mask.astype(NP[real])


def oracle_model(real, grid, mask, a=0.4, b=0.05):
    return CurvatureFlowOracle(oracle_lib(real), grid, mask, a=a, b=b)


@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("masked", [False, True])
def test_curvature_stages_and_ghost_fill_match_oracle(real, masked, device):
    grid, phi, mask = noisy_case(real)
    om = oracle_model(real, grid, mask if masked else None)
    om.do_mask = masked
    # oracle: stage 1, ghost fill, mask; then stage 2 in place
    p = phi.copy()
    om._bc(p)
    dt = 0.9 * om.step(p.copy(), None, 0.0, 1.0)
    want_s1 = np.zeros_like(p)
    om._rhs(p, False)
    o = om.o
    o.lsm3dtvdrk2stage1_(want_s1, grid.gb, p, grid.gb, om.rhs, grid.gb, grid.fb, dt)
    om._bc(want_s1)
    if masked:
        impose_mask(want_s1, mask, want_s1)
    om._rhs(want_s1, False)
    want_nxt = np.zeros_like(p)
    o.lsm3dtvdrk2stage2_(want_nxt, grid.gb, want_s1, grid.gb, p, grid.gb, om.rhs, grid.gb, grid.fb, dt)
    om._bc(want_nxt)
    if masked:
        impose_mask(want_nxt, mask, want_nxt)

    m = CurvatureFlow(grid, real, torch.from_numpy(mask).to(device) if masked else None, device=device)
    phi_d = torch.from_numpy(p).to(device)
    m.curvature_stage(RK1, m.s1, phi_d, None, dt)
    m.fill_ghosts(m.s1)
    torch.cuda.synchronize()
    got_s1 = _np(m.s1)
    m.curvature_stage(TVDRK2_STAGE2, phi_d, m.s1, phi_d, dt)
    m.fill_ghosts(phi_d)
    torch.cuda.synchronize()
    got = _np(phi_d)
    if real == "double":
        assert np.array_equal(got_s1, want_s1)
        assert np.array_equal(got, want_nxt)
    else:
        assert _rel(got_s1, want_s1) < 2e-6
        assert _rel(got, want_nxt) < 2e-6
    if masked:
        assert np.all(got >= mask)


@pytest.mark.parametrize("real", ["double", "float"])
def test_ghost_fill_matches_six_face_sequence(real, device):
    """Random values everywhere: the one-pass ghost kernel equals the six
    LSM3D_SIGNED_LINEAR_EXTRAPOLATION calls (stale ghost values included),
    from phi itself and from a separate source array, with and without mask."""
    rng = np.random.default_rng(8)
    grid = Grid((9, 7, 6), (0.0,) * 3, (1.0,) * 3, 3)
    lib = toolbox(real)
    ora = oracle_lib(real)
    for use_src in (False, True):
        for masked in (False, True):
            a = rng.standard_normal(grid.shape).astype(NP[real])
            src = rng.standard_normal(grid.shape).astype(NP[real])
            mask = rng.standard_normal(grid.shape).astype(NP[real])
            want = a.copy()
            if use_src:
                want[grid.ghost:-grid.ghost, grid.ghost:-grid.ghost, grid.ghost:-grid.ghost] = \
                    src[grid.ghost:-grid.ghost, grid.ghost:-grid.ghost, grid.ghost:-grid.ghost]
            for face in range(6):
                ora.lsm3dsignedlinearextrapolation_(want, grid.gb, grid.fb, face)
            if masked:
                impose_mask(want, mask, want)
            # the kernel writes ghost cells only: the fill box keeps a's values
            want[grid.ghost:-grid.ghost, grid.ghost:-grid.ghost, grid.ghost:-grid.ghost] = \
                a[grid.ghost:-grid.ghost, grid.ghost:-grid.ghost, grid.ghost:-grid.ghost]
            ad = torch.from_numpy(a).to(device)
            sd = torch.from_numpy(src).to(device) if use_src else None
            md = torch.from_numpy(mask).to(device) if masked else None
            from lsmlib_amd.pylsmlib import _ptr, box_c
            gb = box_c(grid.gb)
            lib.fused("lsm3d_amd_extrapolate_ghosts_masked", _ptr(ad), gb, _ptr(sd) if use_src else None, gb,
                      _ptr(md) if masked else None, gb, box_c(grid.fb), box_c((3,) * 6))
            torch.cuda.synchronize()
            assert np.array_equal(_np(ad), want), (use_src, masked)


@pytest.mark.parametrize("real", ["double", "float"])
def test_normal_dt_matches_toolbox_chain(real, device):
    grid, phi, mask = noisy_case(real)
    om = oracle_model(real, grid, mask)
    om._bc(phi)
    om._eno2(phi)
    g = grid
    want = om.o.lsm3dcomputestableconstnormalveldt_(0.4, om.xp, om.yp, om.zp, g.gb, om.xm, om.ym, om.zm, g.gb, g.fb,
                                                    *g.dx, 0.5)
    m = CurvatureFlow(grid, real, None, device=device)
    got = m.normal_dt(torch.from_numpy(phi).to(device))
    if real == "double":
        assert got == want
    else:
        assert got == pytest.approx(want, rel=1e-6)


@pytest.mark.parametrize("real", ["double", "float"])
def test_reinit_eno2_stage_matches_oracle(real, device):
    grid, phi, mask = noisy_case(real)
    om = oracle_model(real, grid, mask)
    om._bc(phi)
    want = phi.copy()
    om.reinitialize(want, 2.5 * grid.dx[0])   # 5 TVD-RK2 iterations, masked
    m = CurvatureFlow(grid, real, torch.from_numpy(mask).to(device), device=device)
    got_d = torch.from_numpy(phi.copy()).to(device)
    n = m.reinitialize(got_d, 2.5 * grid.dx[0])
    torch.cuda.synchronize()
    assert n >= 5
    got = _np(got_d)
    if real == "double":
        assert np.array_equal(got, want)
    else:
        assert _rel(got, want) < 1e-5


@pytest.mark.parametrize("real", ["double", "float"])
def test_pore_throat_run_matches_oracle(real, device):
    """The example's main loop on its default problem for one interval
    (41 steps, then 11 reinitialization iterations) against the oracle of the
    same precision: same step count and statistics; FP64 fields bit for bit,
    FP32 fields within 1e-4 of max|phi| (the float oracle keeps the
    reference float build's double-precision literals, the kernels compute
    in FP32 throughout; the largest differences, ~4e-5, sit at the mask's
    kink on the boundary) and within 1e-5 dx of the FP64 oracle where
    |phi| < 3 dx."""
    g, phi, mask = pore_throat(0.04)
    npdt = NP[real]
    om = CurvatureFlowOracle(oracle_lib(real), g, mask.numpy().astype(npdt))
    p = phi.numpy().astype(npdt)
    want = om.run(p, max_outer=1)
    m = CurvatureFlow(g, real, mask.to(device=device, dtype=DT[real]), device=device)
    pd = phi.to(device=device, dtype=DT[real])
    got = m.run(pd, max_outer=1)
    torch.cuda.synchronize()
    assert got["reinit_iters"] == want["reinit_iters"] and got["steps"] == want["steps"]
    gp = _np(pd)
    if real == "double":
        # the volume is a device sum (FP64, another order): equal to rounding
        assert got["dt_min"] == want["dt_min"] and got["dt_max"] == want["dt_max"]
        for (t1, e1, v1), (t2, e2, v2) in zip(got["history"], want["history"]):
            assert t1 == t2 and e1 == e2 and v1 == pytest.approx(v2, rel=1e-13)
        assert np.array_equal(gp, p)
    else:
        assert got["dt_max"] == pytest.approx(want["dt_max"], rel=1e-5)
        assert _rel(gp, p) < 1e-4
        # against FP64: the volume (the float oracle sums in FP32, the
        # library in FP64) and the field near the interface, |phi| < 3 dx
        p64 = phi.numpy().copy()
        w64 = CurvatureFlowOracle(oracle_lib("double"), g, mask.numpy().copy()).run(p64, max_outer=1)
        assert got["vol_phi"] == pytest.approx(w64["vol_phi"], rel=1e-6)
        near = np.abs(g.interior(p64)) < 3 * g.dx[0]
        err = np.abs(g.interior(gp).astype(np.float64) - g.interior(p64))[near].max()
        assert err < 1e-5 * g.dx[0], err


@pytest.mark.parametrize("real", ["double", "float"])
def test_tilings_give_identical_results(real, device):
    grid, phi, mask = noisy_case(real, n=(70, 13, 23))
    lib = toolbox(real)
    outs = []
    for cfg, kz in ((0, 32), (1, 5), (2, 8), (3, 1), (4, 64), (5, 3), (6, 7)):
        m = CurvatureFlow(grid, real, torch.from_numpy(mask).to(device), device=device, tiling=(cfg, kz))
        pd = torch.from_numpy(phi).to(device)
        m.fill_ghosts(pd, masked=False)
        m.step(pd, 1e-3)
        m.reinitialize(pd, grid.dx[0])
        torch.cuda.synchronize()
        outs.append(_np(pd))
    lib.lib.lsm3d_amd_set_stage_tiling(0, 32)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("real", ["double", "float"])
def test_impose_mask(real, device):
    rng = np.random.default_rng(2)
    a = torch.from_numpy(rng.standard_normal(1000).astype(NP[real])).to(device)
    mk = torch.from_numpy(rng.standard_normal(1000).astype(NP[real])).to(device)
    want = torch.where(mk > a, mk, a)
    out = torch.empty_like(a)
    from lsmlib_amd.pylsmlib import _ptr
    lib = toolbox(real)
    lib.fused("lsm3d_amd_impose_mask", _ptr(out), _ptr(mk), _ptr(a), 1000)
    lib.fused("lsm3d_amd_impose_mask", _ptr(a), _ptr(mk), _ptr(a), 1000)   # in place
    torch.cuda.synchronize()
    assert torch.equal(out, want) and torch.equal(a, want)


def test_stage_rejects_bad_arguments(device):
    grid, phi, mask = noisy_case("double", n=(10, 9, 8))
    m = CurvatureFlow(grid, "double", torch.from_numpy(mask).to(device), device=device)
    pd = torch.from_numpy(phi).to(device)
    from lsmlib_amd._native import LsmError
    with pytest.raises(LsmError):
        m.curvature_stage(32, m.s1, pd, None, 1e-3)                # not an RK2 form
    with pytest.raises(LsmError):
        m.curvature_stage(TVDRK2_STAGE2, m.s1, pd, None, 1e-3)     # stage 2 without phi_cur
    m.raw = None
    with pytest.raises(LsmError):
        m.curvature_stage(RK1, m.s1, pd, None, 1e-3)               # mask without phi_raw
