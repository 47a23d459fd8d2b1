"""Pins the oracle of the curvature model path (CPU only).

1. The new toolbox restatements (central gradients of order 2 and 4, the
   second-order Laplacian, the gradient magnitude, the constant-curvature
   right-hand-side term) are exact where the mathematics makes them exact:
   central differences of order p on polynomials of degree p (order 4: up to
   degree 4), the composed second derivatives and the curvature term on
   general quadratics (every mixed term weighted differently, so a swapped
   index, a lost factor 2 or a wrong sign fails), a sphere's mean curvature
   converging at second order.  The golden-vector and reference-build checks
   of tests/test_oracle.py cover the same routines bit for bit.
2. The composed model step (oracle/steps.py:CurvatureFlowOracle): bit for bit
   the reference build's composition where oracle/_ref is built; a planar
   front moves at exactly the normal speed; the mask constraint holds on
   every cell after every step; a sphere under pure curvature flow shrinks
   as R^2 = R0^2 - 4 b t.
"""
import numpy as np
import pytest

from lsmlib_amd.f77 import box_shape
from lsmlib_amd.pylsmlib import Grid
from lsmlib_amd.workloads import pore_throat
from oracle.oracle import oracle_lib, ref_lib
from oracle.steps import CurvatureFlowOracle

FB = (0, 9, 0, 8, 0, 7)
G = 3
GB = (FB[0] - G, FB[1] + G, FB[2] - G, FB[3] + G, FB[4] - G, FB[5] + G)
DX = (0.1, 0.07, 0.13)


def coords():
    i = np.arange(GB[0], GB[1] + 1)
    j = np.arange(GB[2], GB[3] + 1)
    k = np.arange(GB[4], GB[5] + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return 0.3 + I * DX[0], -0.2 + J * DX[1], 0.1 + K * DX[2]


def interior(a, w=G):
    return a[w:a.shape[0] - w, w:a.shape[1] - w, w:a.shape[2] - w]


def zeros():
    return np.zeros(box_shape(GB))


@pytest.fixture(scope="module")
def ora():
    return oracle_lib("double")


@pytest.mark.parametrize("order,degree", [(2, 2), (4, 4)])
def test_central_gradient_exact_on_polynomials(ora, order, degree):
    X, Y, Z = coords()
    c = [0.4, -1.1, 0.7, 0.9, -0.6]
    p = lambda t: sum(c[m] * t ** m for m in range(degree + 1))  # noqa: E731
    dp = lambda t: sum(m * c[m] * t ** (m - 1) for m in range(1, degree + 1))  # noqa: E731
    phi = p(X) + 2 * p(Y) - p(Z)
    g = [zeros() for _ in range(3)]
    getattr(ora, f"lsm3dcentralgradorder{order}_")(*g, GB, phi, GB, FB, *DX)
    for got, want in zip(g, (dp(X), 2 * dp(Y), -dp(Z))):
        np.testing.assert_allclose(interior(got), interior(want), rtol=0, atol=1e-10)
    if order == 2:  # not exact one degree up
        phi3 = X ** 3
        ora.lsm3dcentralgradorder2_(*g, GB, phi3, GB, FB, *DX)
        assert np.abs(interior(g[0]) - interior(3 * X ** 2)).max() > 1e-4


def test_laplacian_exact_on_quadratics(ora):
    X, Y, Z = coords()
    phi = 1.5 * X ** 2 - 0.5 * Y ** 2 + 2.0 * Z ** 2 + X * Y - 3 * Y * Z + 0.2 * X
    lap = zeros()
    ora.lsm3dlaplacianorder2_(lap, GB, phi, GB, FB, *DX)
    np.testing.assert_allclose(interior(lap), 2 * (1.5 - 0.5 + 2.0), rtol=0, atol=1e-9)


def test_gradient_magnitude_thresholds(ora):
    rng = np.random.default_rng(4)
    g = [rng.standard_normal(box_shape(GB)) for _ in range(3)]
    g[0][0, 0, :] = 5e-5        # square 2.5e-9 < 1e-8: dropped
    g[1][1, :, :] = 0.0
    mag = zeros()
    ora.lsm3dgradientmagnitude_(*g, mag, GB, FB)
    sq = [np.where(a * a < 1e-8, 0.0, a * a) for a in g]
    want = np.sqrt(sq[0] + sq[1] + sq[2])
    want = np.where(want < 1e-8, 0.0, want)
    np.testing.assert_allclose(interior(mag), interior(want), rtol=1e-15)


def _curvature_composed(ora, phi, b=1.0, gb=GB, fb=FB, dx=DX):
    """The example's composition: central gradient of phi on the ghost box
    shrunk by 1, then of phi_x, phi_y, phi_z on the box shrunk by 2, then
    ADD_CONST_CURV over the fill box."""
    z = lambda: np.zeros(box_shape(gb))  # noqa: E731
    s1 = (gb[0] + 1, gb[1] - 1, gb[2] + 1, gb[3] - 1, gb[4] + 1, gb[5] - 1)
    s2 = (gb[0] + 2, gb[1] - 2, gb[2] + 2, gb[3] - 2, gb[4] + 2, gb[5] - 2)
    px, py, pz, pxx, pxy, pxz, pyy, pyz, pzz, rhs = (z() for _ in range(10))
    ora.lsm3dcentralgradorder2_(px, py, pz, gb, phi, gb, s1, *dx)
    ora.lsm3dcentralgradorder2_(pxx, pxy, pxz, gb, px, gb, s2, *dx)
    ora.lsm3dcentralgradorder2_(pxy, pyy, pyz, gb, py, gb, s2, *dx)
    ora.lsm3dcentralgradorder2_(pxz, pyz, pzz, gb, pz, gb, s2, *dx)
    ora.lsm3daddconstcurvtermtolserhs_(rhs, gb, px, py, pz, gb, pxx, pxy, pxz, pyy, pyz, pzz, gb, b, fb)
    return rhs


def test_curvature_term_exact_on_quadratics(ora):
    """Central differences are exact on quadratics, so the composed term is
    N / |grad phi|^2 of the exact derivatives; every Hessian entry differs."""
    X, Y, Z = coords()
    H = np.array([[1.0, 0.6, -0.45], [0.6, 2.2, 0.25], [-0.45, 0.25, -0.8]])
    gvec = np.array([0.9, -0.4, 1.3])
    phi = gvec[0] * X + gvec[1] * Y + gvec[2] * Z + 0.5 * (
        H[0, 0] * X * X + H[1, 1] * Y * Y + H[2, 2] * Z * Z) + H[0, 1] * X * Y + H[0, 2] * X * Z + H[1, 2] * Y * Z
    b = 0.7
    rhs = _curvature_composed(ora, phi, b)
    gx = gvec[0] + H[0, 0] * X + H[0, 1] * Y + H[0, 2] * Z
    gy = gvec[1] + H[0, 1] * X + H[1, 1] * Y + H[1, 2] * Z
    gz = gvec[2] + H[0, 2] * X + H[1, 2] * Y + H[2, 2] * Z
    n = (H[0, 0] * (gy ** 2 + gz ** 2) + H[1, 1] * (gx ** 2 + gz ** 2) + H[2, 2] * (gx ** 2 + gy ** 2)
         - 2 * (H[0, 1] * gx * gy + H[0, 2] * gx * gz + H[1, 2] * gy * gz))
    want = b * n / (gx ** 2 + gy ** 2 + gz ** 2)
    np.testing.assert_allclose(interior(rhs), interior(want), rtol=1e-10, atol=1e-10)


def test_curvature_term_zero_gradient_branch(ora):
    """|grad phi|^2 below zero_tol gives no contribution."""
    phi = np.full(box_shape(GB), 0.25)
    rhs = np.full(box_shape(GB), 3.0)
    assert np.array_equal(interior(_curvature_composed(ora, phi) + rhs), interior(rhs))


def test_sphere_mean_curvature_converges_second_order(ora):
    """phi = |x| - R: b kappa |grad phi| = 2b/r; error at r in [0.45, 0.55]
    drops ~4x when h halves."""
    errs = []
    for n in (24, 48):
        h = 2.0 / n
        fb = (0, n - 1, 0, n - 1, 0, n - 1)
        gb = (-3, n + 2, -3, n + 2, -3, n + 2)
        ax = -1 + (np.arange(-3, n + 3) + 0.5) * h
        Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
        r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
        rhs = _curvature_composed(ora, r - 0.5, 1.0, gb, fb, (h, h, h))
        sel = (r > 0.45) & (r < 0.55)
        errs.append(np.abs(rhs[sel] - 2 / r[sel]).max())
    assert errs[1] < errs[0] / 3.0, errs


# ---------------------------------------------------------------------------
# the composed model step
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("real", ["double", "float"])
def test_model_oracle_matches_reference_build(real):
    ref = ref_lib(real)
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    g, phi, mask = pore_throat(0.04)
    dt = np.float64 if real == "double" else np.float32
    runs = []
    for lib in (oracle_lib(real), ref):
        p, m = phi.numpy().astype(dt), mask.numpy().astype(dt)
        st = CurvatureFlowOracle(lib, g, m).run(p, max_outer=1, max_steps=12)
        runs.append((p, st))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1]["history"] == runs[1][1]["history"]


def _plane_grid(n=12):
    h = 1.0 / n
    return Grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3, spacing=(h, h, h))


def test_planar_front_moves_at_normal_speed():
    """phi = x - c: zero curvature, |grad phi| = 1 exactly under ENO2 and the
    signed linear fill, so a TVD-RK2 step subtracts a dt everywhere."""
    ora = oracle_lib("double")
    grid = _plane_grid()
    gb = grid.gb
    ax = [(np.arange(gb[2 * d], gb[2 * d + 1] + 1) + 0.5) / grid.n[d] for d in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    phi = X - 0.37
    model = CurvatureFlowOracle(ora, grid, None, a=0.4, b=0.05, do_mask=False)
    dt = model.step(phi.copy(), None, 0.0, 0.1)   # the stable dt
    p = X - 0.37
    model.step(p, dt)
    np.testing.assert_allclose(grid.interior(p), grid.interior(X - 0.37 - 0.4 * dt), rtol=0, atol=1e-13)


def test_mask_holds_after_every_step():
    ora = oracle_lib("double")
    g, phi, mask = pore_throat(0.04)
    p, m = phi.numpy().copy(), mask.numpy().copy()
    model = CurvatureFlowOracle(ora, g, m)
    dt_sub = 0.0
    for _ in range(4):
        dt_sub += model.step(p, None, dt_sub, 0.1)
        assert np.all(p >= m)
    model.reinitialize(p, 5 * g.dx[0])
    assert np.all(p >= m)


def test_sphere_shrinks_under_mean_curvature_flow():
    """a = 0, b = 1: dR/dt = -2/R, so R^2 = R0^2 - 4t (radius from the
    enclosed volume, 5% band for the O(h^2) scheme on a 40^3 grid)."""
    ora = oracle_lib("double")
    n = 40
    h = 2.0 / n
    grid = Grid((n, n, n), (-1.0,) * 3, (1.0,) * 3, 3, spacing=(h, h, h))
    gb = grid.gb
    ax = -1 + (np.arange(gb[0], gb[1] + 1) + 0.5) * h
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    r0 = 0.7
    phi = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - r0
    model = CurvatureFlowOracle(ora, grid, None, a=0.0, b=1.0, do_mask=False)
    t = 0.0
    for _ in range(40):
        t += model.step(phi, None, 0.0, 1.0)
    vol = ora.lsm3dvolumeregionphilessthanzero_(phi, grid.gb, grid.fb, h, h, h, 1.5 * h)
    r = (3 * vol / (4 * np.pi)) ** (1 / 3)
    assert r == pytest.approx(np.sqrt(r0 ** 2 - 4 * t), rel=0.05), (r, t)
    assert r < r0 - 0.02
