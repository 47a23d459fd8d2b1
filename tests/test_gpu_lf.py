"""Lax-Friedrichs Hamiltonians on the GPU (no counterpart in the reference).

The reference's level-set right-hand sides are upwind (advection) or
Godunov (normal speed) only, so these routines have no reference output to
match.  They are pinned instead by what the mathematics fixes, each against
routines that are themselves pinned to the reference build:

* local Lax-Friedrichs (alpha_a = |v_a|) equals the upwind advection term
  LSM3D_UPWIND_HJ_WENO5 + LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS bit for bit in
  FP64 wherever |v_a| >= LSMLIB_ZERO_TOL (the oracle's chain);
* for any alpha the term differs from that upwind term by exactly the
  numerical viscosity sum_a 0.5 (alpha_a - |v_a|)(phi_a+ - phi_a-), with the
  plus/minus derivatives of the oracle's LSM3D_HJ_WENO5;
* the normal-speed Hamiltonian on a planar signed distance (phi+ = phi-,
  |grad phi| = 1) gives rhs = -vn for local and global alpha;
* the fused stage equals the drop-in chain LSM3D_HJ_WENO5 ->
  LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS -> lsm3d_amd_add_lf_advection_term ->
  LSM3D_TVD_RK3_STAGE* on the device, bit for bit in FP64;
* global Lax-Friedrichs WENO5 / TVD-RK3 converges on a smooth translated
  field (observed order >= 2.5 between 24, 48 and 96 cells per axis).
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from lsmlib_amd import toolbox
from lsmlib_amd.pylsmlib import RK1, TVDRK3_STAGE2, TVDRK3_STAGE3, BC_PERIODIC, Grid, Stepper, box_c
from oracle.oracle import oracle_lib

pytestmark = pytest.mark.gpu


def _field(shape, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(shape)


def _case(seed=1, shape=(17, 19, 22)):
    """A smooth-plus-noise phi on a 16 x 13 x 11 fill box with 3 ghosts on a
    shifted box, and a velocity with |v| in [1e-3, 2] and random signs."""
    g = Grid((shape[2] - 6, shape[1] - 6, shape[0] - 6), (0.0, 0.0, 0.0), (1.0, 0.9, 0.8), 3)
    X, Y, Z = (a.numpy() for a in g.centers(torch.float64))
    phi = np.sin(3 * X) * np.cos(2 * Y) + Z ** 2 - 0.3 + 0.01 * _field(X.shape, seed)
    rng = np.random.default_rng(seed + 7)
    vel = tuple(rng.choice([-1.0, 1.0], size=X.shape) * rng.uniform(1e-3, 2.0, size=X.shape) for _ in range(3))
    return g, np.ascontiguousarray(phi), tuple(np.ascontiguousarray(v) for v in vel)


def _oracle_pm(g, phi):
    o = oracle_lib("double")
    z = [np.zeros(g.shape) for _ in range(7)]
    o.lsm3dhjweno5_(z[0], z[1], z[2], g.gb, z[3], z[4], z[5], g.gb, phi, g.gb, z[6], g.gb, g.fb, *g.dx)
    return z[:3], z[3:6]


def _oracle_upwind_rhs(g, phi, vel):
    o = oracle_lib("double")
    gx, gy, gz, d1, rhs = (np.zeros(g.shape) for _ in range(5))
    o.lsm3dupwindhjweno5_(gx, gy, gz, g.gb, phi, g.gb, vel[0], vel[1], vel[2], g.gb, d1, g.gb, g.fb, *g.dx)
    o.lsm3dzerooutlevelseteqnrhs_(rhs, g.gb)
    o.lsm3daddadvectiontermtolserhs_(rhs, g.gb, gx, gy, gz, g.gb, vel[0], vel[1], vel[2], g.gb, g.fb)
    return rhs


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dtype, device="cuda")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _lf_advection(g, plus, minus, vel, alpha=None, real="double"):
    lib = toolbox(real)
    dt = torch.float64 if real == "double" else torch.float32
    rhs = torch.zeros(g.shape, dtype=dt, device="cuda")
    P = [_dev(a, dt) for a in plus]
    M = [_dev(a, dt) for a in minus]
    V = [_dev(a, dt) for a in vel]
    al = (lib._ct * 3)(*alpha) if alpha is not None else None
    gb, fb = box_c(g.gb), box_c(g.fb)
    lib.fused("lsm3d_amd_add_lf_advection_term", _p(rhs), gb, *map(_p, P), gb, *map(_p, M), gb, *map(_p, V), gb, fb,
              1 if alpha is not None else 0, al)
    torch.cuda.synchronize()
    return rhs.cpu().numpy()


def test_local_lf_equals_reference_upwind_bitwise():
    g, phi, vel = _case()
    plus, minus = _oracle_pm(g, phi)
    got = _lf_advection(g, plus, minus, vel)
    want = _oracle_upwind_rhs(g, phi, vel)
    fb = (slice(3, -3),) * 3
    assert np.array_equal(got[fb], want[fb]), float(np.abs(got[fb] - want[fb]).max())


def test_global_lf_adds_exactly_the_numerical_viscosity():
    g, phi, vel = _case(seed=3)
    plus, minus = _oracle_pm(g, phi)
    fb = (slice(3, -3),) * 3
    alpha = tuple(1.25 * float(np.abs(v[fb]).max()) for v in vel)
    got = _lf_advection(g, plus, minus, vel, alpha)
    up = _oracle_upwind_rhs(g, phi, vel)
    visc = sum(0.5 * (alpha[a] - np.abs(vel[a])) * (plus[a] - minus[a]) for a in range(3))
    want = up + visc
    scale = np.abs(up[fb]).max()
    assert np.abs(got[fb] - want[fb]).max() <= 1e-13 * scale
    # the viscosity is what makes it differ from upwind (not a trivially equal case)
    assert np.abs(got[fb] - up[fb]).max() > 1e-6 * scale


@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("mode", ["local", "global"])
def test_lf_normal_speed_on_a_planar_distance(real, mode):
    g = Grid((14, 12, 10), (-0.5, -0.4, -0.3), (0.5, 0.4, 0.3), 3)
    X, Y, Z = (a.numpy() for a in g.centers(torch.float64))
    n = np.array([2.0, -1.0, 2.0]) / 3.0  # unit normal, exact in binary to a few ulps
    phi = n[0] * X + n[1] * Y + n[2] * Z - 0.05
    plus, minus = _oracle_pm(g, np.ascontiguousarray(phi))
    vn = np.random.default_rng(5).uniform(-2.0, 2.0, size=X.shape)
    lib = toolbox(real)
    dt = torch.float64 if real == "double" else torch.float32
    rhs = torch.zeros(g.shape, dtype=dt, device="cuda")
    P = [_dev(a, dt) for a in plus]
    M = [_dev(a, dt) for a in minus]
    V = _dev(vn, dt)
    gb, fb = box_c(g.gb), box_c(g.fb)
    alpha = float(np.abs(vn).max()) * 1.1 if mode == "global" else 0.0
    lib.fused("lsm3d_amd_add_lf_normal_vel_term", _p(rhs), gb, *map(_p, P), gb, *map(_p, M), gb, _p(V), gb, 0.0,
              fb, 1 if mode == "global" else 0, alpha)
    torch.cuda.synchronize()
    got = rhs.cpu().numpy()[(slice(3, -3),) * 3]
    want = -vn[(slice(3, -3),) * 3]
    tol = 1e-13 if real == "double" else 2e-6
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


def _periodic_case(n, real, seed=2):
    g = Grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
    X, Y, Z = g.centers(torch.float64)
    phi = torch.sin(2 * math.pi * X) * torch.cos(2 * math.pi * Y) + 0.5 * torch.cos(2 * math.pi * Z) - 0.1
    rng = np.random.default_rng(seed)
    vel = tuple(torch.from_numpy(rng.choice([-1.0, 1.0], size=X.shape) * rng.uniform(1e-2, 1.5, size=X.shape))
                for _ in range(3))
    dt = torch.float64 if real == "double" else torch.float32
    return g, phi.to(dt).contiguous(), tuple(v.to(dt).contiguous() for v in vel)


@pytest.mark.parametrize("alpha", [None, (2.0, 2.0, 2.0)])
def test_fused_lf_stage_equals_the_drop_in_chain(alpha):
    real = "double"
    g, phi, vel = _periodic_case(20, real)
    st = Stepper(g, real, (BC_PERIODIC,) * 6, ghosts="fill")
    dev = torch.device("cuda")
    p = phi.to(dev).clone()
    V = tuple(v.to(dev) for v in vel)
    st.fill_ghosts(p)
    s1, s2 = st.empty(), st.empty()
    dt_ = 0.2 * g.dx[0]
    ref = p.clone()
    st.advect_lf_rk3_step(p, s1, s2, V, dt_, alpha)
    # the same step from the device drop-in routines
    lib = toolbox(real)
    gb, fbx = box_c(g.gb), box_c(g.fb)
    pm = [torch.zeros(g.shape, dtype=torch.float64, device=dev) for _ in range(7)]
    rhs = torch.zeros_like(ref)
    al = (lib._ct * 3)(*alpha) if alpha is not None else None

    def lf_rhs(ps):
        lib.lsm3dhjweno5_(pm[0], pm[1], pm[2], g.gb, pm[3], pm[4], pm[5], g.gb, ps, g.gb, pm[6], g.gb, g.fb, *g.dx)
        lib.lsm3dzerooutlevelseteqnrhs_(rhs, g.gb)
        lib.fused("lsm3d_amd_add_lf_advection_term", _p(rhs), gb, *map(_p, pm[:3]), gb, *map(_p, pm[3:6]), gb,
                  *map(_p, V), gb, fbx, 1 if alpha is not None else 0, al)
        return rhs

    u1, u2, un = ref.clone(), ref.clone(), ref.clone()
    lib.lsm3dtvdrk3stage1_(u1, g.gb, ref, g.gb, lf_rhs(ref), g.gb, g.fb, dt_)
    st.fill_ghosts(u1)
    lib.lsm3dtvdrk3stage2_(u2, g.gb, u1, g.gb, ref, g.gb, lf_rhs(u1), g.gb, g.fb, dt_)
    st.fill_ghosts(u2)
    lib.lsm3dtvdrk3stage3_(un, g.gb, u2, g.gb, ref, g.gb, lf_rhs(u2), g.gb, g.fb, dt_)
    torch.cuda.synchronize()
    a, b = g.interior(p).cpu().numpy(), g.interior(un).cpu().numpy()
    assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_fused_lf_fp32_close_to_fp64():
    out = {}
    for real in ("double", "float"):
        g, phi, vel = _periodic_case(24, real)
        st = Stepper(g, real, (BC_PERIODIC,) * 6)
        p = phi.cuda().clone()
        V = tuple(v.cuda() for v in vel)
        st.fill_ghosts(p)
        s1, s2 = st.empty(), st.empty()
        for _ in range(3):
            st.advect_lf_rk3_step(p, s1, s2, V, 0.2 * g.dx[0])
        torch.cuda.synchronize()
        out[real] = g.interior(p).double().cpu().numpy()
    assert np.abs(out["float"] - out["double"]).max() <= 1e-5 * np.abs(out["double"]).max()


def test_global_lf_converges_on_a_smooth_translated_field():
    """phi0 = sin 2pi x cos 2pi y + 0.5 cos 2pi z on the periodic unit cube,
    constant V = (1, 0.5, 0.25), global alpha = 1.5 |V_a|, t = 0.05 with
    dt = 0.4 dx: the error against phi0(x - V t) falls at order >= 2.5."""
    errs = []
    V = (1.0, 0.5, 0.25)
    T = 0.05
    for n in (24, 48, 96):
        g = Grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
        X, Y, Z = g.centers(torch.float64, "cuda")

        def f(x, y, z):
            return torch.sin(2 * math.pi * x) * torch.cos(2 * math.pi * y) + 0.5 * torch.cos(2 * math.pi * z)

        p = f(X, Y, Z).contiguous()
        vel = tuple(torch.full_like(X, v).contiguous() for v in V)
        st = Stepper(g, "double", (BC_PERIODIC,) * 6)
        st.fill_ghosts(p)
        s1, s2 = st.empty(), st.empty()
        steps = int(round(T / (0.4 * g.dx[0])))
        dt = T / steps
        for _ in range(steps):
            st.advect_lf_rk3_step(p, s1, s2, vel, dt, alpha=tuple(1.5 * abs(v) for v in V))
        exact = f(X - V[0] * T, Y - V[1] * T, Z - V[2] * T)
        errs.append(float((g.interior(p) - g.interior(exact)).abs().max()))
        del X, Y, Z, p, exact, vel, s1, s2
    rates = [math.log2(errs[q] / errs[q + 1]) for q in range(2)]
    assert min(rates) >= 2.5, (errs, rates)
