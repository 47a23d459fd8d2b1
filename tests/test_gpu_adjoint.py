"""GPU checks of the reverse-mode WENO5 / TVD-RK advection path.

The reference has no adjoint, so the stage adjoint is pinned to the
derivative of the reference's forward stage: central finite differences of
the oracle's LSM3D_UPWIND_HJ_WENO5 -> ADD_ADVECTION_TERM -> TVD_RK stage
chain (oracle/steps.py, bit-exact with the reference build) in FP64, tested
through the dot-product identity <ubar, dS[delta]> = <S^T ubar, delta> along
random directions for phi_stage, phi_cur and each velocity component.  The
whole checkpointed TVD-RK3 run is checked the same way through the
Heaviside mismatch loss (FD of the GPU FP64 forward), for checkpoint
invariance, and FP32 against FP64.  The loss kernel is checked against the
oracle's lsmHeaviside (lsm_calculus_toolbox.f:25-54).
"""
import ctypes

import numpy as np
import pytest
import torch

from lsmlib_amd import LsmError
from lsmlib_amd.adjoint import AdvectionRK3, advect_weno5_rk3, heaviside_mismatch
from lsmlib_amd.pylsmlib import BC_PERIODIC, RK1, TVDRK2_STAGE2, TVDRK3_STAGE2, TVDRK3_STAGE3, Grid, Stepper
from lsmlib_amd.workloads import adjoint_case
from oracle.oracle import oracle_lib
from oracle.steps import OracleStepper, fill_ghosts

pytestmark = pytest.mark.gpu

PERIODIC = (BC_PERIODIC,) * 6


def _dot(a, b):
    return float((a.astype(np.float64) * b.astype(np.float64)).sum())


def _interior(grid, a):
    g = grid.ghost
    return a[g:-g, g:-g, g:-g]


class OracleStage:
    """One forward TVD-RK stage of the reference chain on the CPU, periodic."""

    def __init__(self, grid):
        self.ora = oracle_lib("double")
        self.grid = grid
        self.st = OracleStepper(self.ora, grid, PERIODIC)

    def __call__(self, form, us, uc, vel, dt):
        o, g = self.ora, self.grid
        us = us.copy()
        fill_ghosts(o, us, g, PERIODIC)
        rhs = self.st.advect_rhs(us, vel).copy()
        out = np.zeros_like(us)
        if form == RK1:
            o.lsm3dtvdrk3stage1_(out, g.gb, us, g.gb, rhs, g.gb, g.fb, dt)
        elif form == TVDRK2_STAGE2:
            o.lsm3dtvdrk2stage2_(out, g.gb, us, g.gb, uc, g.gb, rhs, g.gb, g.fb, dt)
        elif form == TVDRK3_STAGE2:
            o.lsm3dtvdrk3stage2_(out, g.gb, us, g.gb, uc, g.gb, rhs, g.gb, g.fb, dt)
        else:
            o.lsm3dtvdrk3stage3_(out, g.gb, us, g.gb, uc, g.gb, rhs, g.gb, g.fb, dt)
        return out


def _fields(grid, seed):
    rng = np.random.default_rng(seed)
    shp = grid.shape
    us = rng.normal(size=shp)
    uc = rng.normal(size=shp)
    vel = []
    for _ in range(3):
        v = rng.normal(size=shp)
        v = np.where(np.abs(v) < 1e-2, 0.5, v)  # keep every cell off the upwind switch
        vel.append(v)
    ubar = np.zeros(shp)
    _interior(grid, ubar)[...] = rng.normal(size=_interior(grid, ubar).shape)
    return us, uc, vel, ubar, rng


@pytest.mark.parametrize("form", [RK1, TVDRK2_STAGE2, TVDRK3_STAGE2, TVDRK3_STAGE3])
def test_stage_adjoint_matches_reference_derivative(device, form):
    grid = Grid((12, 14, 16), (0.0, 0.0, 0.0), (1.2, 1.4, 1.6), 3)
    us, uc, vel, ubar, rng = _fields(grid, 100 + form)
    dt = 0.01
    stage = OracleStage(grid)
    runner = AdvectionRK3(grid, "double", device)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    z = lambda: torch.zeros(grid.shape, dtype=torch.float64, device=device)
    us_d = T(us)
    Stepper(grid, "double", PERIODIC, ghosts="images").fill_ghosts(us_d)
    adj_s, adj_c, vbar = z(), z(), [z(), z(), z()]
    runner.stage_adjoint(form, T(ubar), adj_s, adj_c if form != RK1 else None, 0, us_d,
                         tuple(T(v) for v in vel), vbar, dt)
    torch.cuda.synchronize()
    adj_s, adj_c = adj_s.cpu().numpy(), adj_c.cpu().numpy()
    vbar = [v.cpu().numpy() for v in vbar]

    def fd(fun, h=1e-5):
        return (_dot(ubar, fun(h)) - _dot(ubar, fun(-h))) / (2 * h)

    uc_in = uc if form != RK1 else None
    for trial in range(3):
        d = np.zeros(grid.shape)
        _interior(grid, d)[...] = rng.normal(size=_interior(grid, d).shape)
        want = fd(lambda h: stage(form, us + h * d, uc_in, vel, dt))
        got = _dot(_interior(grid, adj_s), _interior(grid, d))
        assert abs(got - want) <= 1e-7 * abs(want), ("phi_stage", trial, got, want)
        if form != RK1:
            want = fd(lambda h: stage(form, us, uc + h * d, vel, dt))
            got = _dot(_interior(grid, adj_c), _interior(grid, d))
            assert abs(got - want) <= 1e-9 * abs(want), ("phi_cur", trial, got, want)
        for a in range(3):
            def moved(h, a=a):
                v = list(vel)
                v[a] = vel[a] + h * d
                return stage(form, us, uc_in, v, dt)
            want = fd(moved)
            got = _dot(_interior(grid, vbar[a]), _interior(grid, d))
            assert abs(got - want) <= 1e-7 * abs(want), ("vel", a, trial, got, want)
    # ghost cells of the outputs are left untouched
    assert not np.any(adj_s - np.pad(_interior(grid, adj_s), 3))


def test_stage_adjoint_accumulate_and_fp32(device):
    grid = Grid((10, 12, 14), (0.0, 0.0, 0.0), (1.0, 1.2, 1.4), 3)
    us, uc, vel, ubar, _ = _fields(grid, 7)
    res = {}
    for real, dt_ in (("double", torch.float64), ("float", torch.float32)):
        r = AdvectionRK3(grid, real, device)
        T = lambda a: torch.tensor(a, dtype=dt_, device=device)
        us_d = T(us)
        Stepper(grid, real, PERIODIC, ghosts="images").fill_ghosts(us_d)
        base_s, base_c = T(uc), T(us)
        s, c = base_s.clone(), base_c.clone()
        r.stage_adjoint(TVDRK3_STAGE2, T(ubar), s, c, 3, us_d, tuple(T(v) for v in vel), None, 0.02)
        s0, c0 = torch.zeros_like(s), torch.zeros_like(c)
        r.stage_adjoint(TVDRK3_STAGE2, T(ubar), s0, c0, 0, us_d, tuple(T(v) for v in vel), None, 0.02)
        torch.cuda.synchronize()
        inner = lambda a: _interior(grid, a.double().cpu().numpy())
        assert np.allclose(inner(s), inner(base_s) + inner(s0), rtol=0, atol=1e-12 if real == "double" else 1e-5)
        assert np.allclose(inner(c), inner(base_c) + inner(c0), rtol=0, atol=1e-12 if real == "double" else 1e-5)
        res[real] = inner(s0)
    err = np.abs(res["float"] - res["double"]).max() / np.abs(res["double"]).max()
    assert err < 1e-4, err


def test_stage_adjoint_rejects_bad_calls(device):
    grid = Grid((8, 8, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
    r = AdvectionRK3(grid, "double", device)
    z = lambda: torch.zeros(grid.shape, dtype=torch.float64, device=device)
    vel = (z(), z(), z())
    with pytest.raises(LsmError, match="periodic"):
        r.lib.fused("lsm3d_amd_advect_stage_adjoint", RK1, *_stage_args(r, z(), z(), vel, periodic=5))
    with pytest.raises(LsmError, match="Runge-Kutta"):
        r.lib.fused("lsm3d_amd_advect_stage_adjoint", 7, *_stage_args(r, z(), z(), vel))
    with pytest.raises(LsmError, match="device memory"):
        host = torch.zeros(grid.shape, dtype=torch.float64)
        r.lib.fused("lsm3d_amd_advect_stage_adjoint", RK1, *_stage_args(r, host, z(), vel))
    with pytest.raises(ValueError):
        r.run(torch.zeros((3, 3, 3), dtype=torch.float64, device=device), vel, 1, 0.1)
    small = Grid((5, 8, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
    rs = AdvectionRK3(small, "double", device)
    zs = lambda: torch.zeros(small.shape, dtype=torch.float64, device=device)
    with pytest.raises(LsmError, match="6 cells"):
        rs.lib.fused("lsm3d_amd_advect_stage_adjoint", RK1, *_stage_args(rs, zs(), zs(), (zs(), zs(), zs())))


def _stage_args(r, ubar, adj, vel, periodic=7):
    gb = r._gb
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    return (P(ubar), gb, P(adj), gb, None, gb, 0, P(ubar), gb, P(vel[0]), P(vel[1]), P(vel[2]), gb,
            None, None, None, gb, P(r.work), r._fb, 0.1, 0.1, 0.1, 0.01, periodic)


def _case(n, real, device):
    dtype = torch.float64 if real == "double" else torch.float32
    grid, phi0, fv, _ = adjoint_case(n, torch.float64, device)
    return grid, phi0, fv, dtype


def _loss_fn(grid, runner, fv, target, nsteps, dt, dtype):
    eps = 1.5 * grid.dx[0]

    def loss(phi0, c):
        vel = fv.velocity(c)
        vel = tuple(v.to(dtype).contiguous() for v in vel)
        phi = advect_weno5_rk3(phi0.to(dtype), vel, runner, nsteps, dt)
        return heaviside_mismatch(phi, target, grid, eps)
    return loss


def _target(grid, runner, fv, phi0, nsteps, dt, dtype, seed=9):
    rng = np.random.default_rng(seed)
    delta = torch.tensor(rng.normal(0.0, 0.02, size=fv.c0.shape), dtype=torch.float64, device=fv.device)
    with torch.no_grad():
        vel = tuple(v.to(dtype).contiguous() for v in fv.velocity(fv.c0.to(fv.device) + delta))
        phi, _ = runner.run(phi0.to(dtype), vel, nsteps, dt, keep_checkpoints=False)
    return phi


def test_adjoint_gradient_matches_finite_differences(device):
    n, nsteps, dt = 24, 6, 0.02
    grid, phi0, fv, dtype = _case(n, "double", device)
    runner = AdvectionRK3(grid, "double", device, checkpoint_every=4)  # two segments, one ragged
    target = _target(grid, runner, fv, phi0, nsteps, dt, dtype)
    loss = _loss_fn(grid, runner, fv, target, nsteps, dt, dtype)
    c = fv.c0.to(device).clone().requires_grad_(True)
    p0 = phi0.clone().requires_grad_(True)
    L = loss(p0, c)
    L.backward()
    gc, gp = c.grad.detach(), p0.grad.detach()
    assert float(L.detach()) > 0
    rng = np.random.default_rng(1)
    # gradient is zero on ghost cells (phi0's ghosts are refilled from the interior)
    g = grid.ghost
    assert float(gp.abs().sum()) == pytest.approx(float(gp[g:-g, g:-g, g:-g].abs().sum()))
    worst = 0.0
    with torch.no_grad():
        for trial in range(4):
            e = torch.tensor(rng.normal(size=gc.shape), dtype=torch.float64, device=device)
            e /= e.norm()
            h = 1e-5
            fd = (float(loss(phi0, fv.c0.to(device) + h * e)) - float(loss(phi0, fv.c0.to(device) - h * e))) / (2 * h)
            ad = float((gc * e).sum())
            worst = max(worst, abs(fd - ad) / abs(ad))
        for trial in range(2):
            X, Y, Z = grid.centers(torch.float64, device)
            w = torch.tensor(rng.uniform(1, 3, size=3))
            e = torch.sin(np.pi * (float(w[0]) * X + float(w[1]) * Y) + float(w[2]) * Z) * torch.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 0.2)
            h = 1e-5
            fd = (float(loss(phi0 + h * e, fv.c0.to(device))) - float(loss(phi0 - h * e, fv.c0.to(device)))) / (2 * h)
            ad = float((gp * e).sum())
            worst = max(worst, abs(fd - ad) / abs(ad))
    assert worst < 1e-5, worst


def test_checkpoint_interval_does_not_change_the_gradient(device):
    n, nsteps, dt = 16, 5, 0.03
    grid, phi0, fv, dtype = _case(n, "double", device)
    out = {}
    for every in (1, 2, 5, 8):
        runner = AdvectionRK3(grid, "double", device, checkpoint_every=every)
        target = _target(grid, runner, fv, phi0, nsteps, dt, dtype)
        c = fv.c0.to(device).clone().requires_grad_(True)
        p0 = phi0.clone().requires_grad_(True)
        L = _loss_fn(grid, runner, fv, target, nsteps, dt, dtype)(p0, c)
        L.backward()
        out[every] = (float(L.detach()), c.grad.clone(), p0.grad.clone())
    for every in (2, 5, 8):
        assert out[every][0] == out[1][0]
        assert torch.equal(out[every][1], out[1][1])
        assert torch.equal(out[every][2], out[1][2])


def test_forward_matches_stepper(device):
    n, nsteps, dt = 16, 4, 0.03
    grid, phi0, fv, dtype = _case(n, "double", device)
    vel = tuple(v.contiguous() for v in fv.velocity(fv.c0.to(device)))
    runner = AdvectionRK3(grid, "double", device, checkpoint_every=3)
    phi = advect_weno5_rk3(phi0, vel, runner, nsteps, dt)
    st = Stepper(grid, "double", PERIODIC, ghosts="fill")
    ref = phi0.clone()
    st.fill_ghosts(ref)
    s1, s2 = st.empty(), st.empty()
    for _ in range(nsteps):
        st.advect_rk3_step(ref, s1, s2, vel, dt)
    torch.cuda.synchronize()
    assert torch.equal(phi, ref)


def test_fp32_gradient_against_fp64(device):
    """The FP32 gradient against the FP64 one on the 128^3, 40-step gradient
    workload, with both runs on the double build's upwind switch (1e-11).
    dL/dc: within 1e-4 (normwise).  dL/dphi0: the FP32 adjoint adds at most
    2 % to what FP32 storage of the forward states already costs (FP64 stage
    adjoints over the same FP32 states), and that storage cost is within
    1.5e-4.  The float build's own switch (1e-5) drops the V_a gradient on
    cells with 1e-11 <= |V_a| < 1e-5, which shows in dL/dc."""
    from lsmlib_amd.precision import gradient_precision_split
    sp = gradient_precision_split(128, 40, 8, device)
    assert sp["fp32"]["c"] <= 1e-4, sp
    assert sp["fp64_adjoint_fp32_states"]["c"] <= 1e-6, sp
    assert sp["fp32"]["phi0"] <= 1.02 * sp["fp64_adjoint_fp32_states"]["phi0"], sp
    assert sp["fp64_adjoint_fp32_states"]["phi0"] <= 1.5e-4, sp
    assert sp["fp64_fp32_inputs"]["phi0"] <= sp["fp64_adjoint_fp32_states"]["phi0"], sp
    assert sp["fp32_switch_1e-5"]["c"] > 10 * sp["fp32"]["c"], sp


def test_fp32_runner_gradient_small_grid(device):
    """The autograd entry in FP32 (switch 1e-11) against FP64 on a 32^3 grid."""
    n, nsteps, dt = 32, 8, 0.01
    grads = {}
    for real in ("double", "float"):
        grid, phi0, fv, dtype = _case(n, real, device)
        runner = AdvectionRK3(grid, real, device, checkpoint_every=8, zero_tol=1e-11)
        target = _target(grid, runner, fv, phi0, nsteps, dt, dtype)
        c = fv.c0.to(device).clone().requires_grad_(True)
        p0 = phi0.to(dtype).clone().requires_grad_(True)
        L = _loss_fn(grid, runner, fv, target, nsteps, dt, dtype)(p0, c)
        L.backward()
        grads[real] = (c.grad.double(), p0.grad.double())
    ec = float((grads["float"][0] - grads["double"][0]).norm() / grads["double"][0].norm())
    ep = float((grads["float"][1] - grads["double"][1]).norm() / grads["double"][1].norm())
    assert ec < 1e-4 and ep < 1e-3, (ec, ep)


def test_reverse_with_fp64_adjoint_runner_equals_fp64_reverse(device):
    """reverse(adjoint_runner=FP64 runner) over FP64 forward states equals the
    FP64 runner's own reverse bit for bit (the cast path is the identity)."""
    n, nsteps, dt = 16, 5, 0.02
    grid, phi0, fv, dtype = _case(n, "double", device)
    r1 = AdvectionRK3(grid, "double", device, checkpoint_every=2)
    r2 = AdvectionRK3(grid, "double", device, checkpoint_every=2)
    vel = tuple(v.contiguous() for v in fv.velocity(fv.c0.to(device)))
    phi, ck = r1.run(phi0, vel, nsteps, dt)
    g = torch.randn(grid.shape, dtype=torch.float64, device=device)
    a = r1.reverse(ck, vel, nsteps, dt, g)
    b = r1.reverse(ck, vel, nsteps, dt, g, adjoint_runner=r2)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0]) and all(torch.equal(x, y) for x, y in zip(a[1], b[1]))


def _oracle_heaviside(real, eps):
    """Vectorised oracle lsmHeaviside (lsm_calculus_toolbox.f:25-54)."""
    fn = oracle_lib(real).lib.oracle_lsmheaviside_
    fn.restype = None
    ct = ctypes.c_double if real == "double" else ctypes.c_float

    def one(x):
        out = ct()
        fn(ctypes.byref(out), ctypes.byref(ct(x)), ctypes.byref(ct(eps)))
        return out.value
    return np.vectorize(one)


@pytest.mark.parametrize("real", ["double", "float"])
def test_heaviside_mismatch_against_oracle(device, real):
    grid = Grid((9, 10, 11), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
    eps = 0.2
    rng = np.random.default_rng(4)
    np_t = np.float64 if real == "double" else np.float32
    phi = rng.uniform(-0.4, 0.4, size=grid.shape).astype(np_t)
    tgt = rng.uniform(-0.4, 0.4, size=grid.shape).astype(np_t)
    phi.flat[:6] = np.array([eps, -eps, 0.0, 2 * eps, -2 * eps, eps / 2], dtype=np_t)
    dtype = torch.float64 if real == "double" else torch.float32
    p = torch.tensor(phi, device=device, requires_grad=True)
    L = heaviside_mismatch(p, torch.tensor(tgt, device=device), grid, eps)
    L.backward()
    H = _oracle_heaviside(real, eps)
    want = (((H(_interior(grid, phi)) - H(_interior(grid, tgt))) ** 2)).sum()
    tol = 1e-13 if real == "double" else 1e-5
    assert abs(float(L) - want) <= tol * want
    grad = p.grad.double().cpu().numpy()
    assert not np.any(grad - np.pad(_interior(grid, grad), 3))
    # d loss / d phi against central differences of the oracle's H, away from the kinks at +-eps
    x = _interior(grid, phi).astype(np.float64).ravel()[:200]
    t = _interior(grid, tgt).astype(np.float64).ravel()[:200]
    keep = np.abs(np.abs(x) - eps) > 1e-3
    h = 1e-6 if real == "double" else 1e-3
    H64 = _oracle_heaviside("double", eps)
    fd = ((H64(x + h) - H64(t)) ** 2 - (H64(x - h) - H64(t)) ** 2) / (2 * h)
    got = _interior(grid, grad).ravel()[:200]
    scale = np.abs(fd).max()
    tol = 1e-7 if real == "double" else 2e-3
    assert np.abs(got[keep] - fd[keep]).max() <= tol * scale
