"""Pins the field-extension path of the oracle (CPU only).

The toolbox routines on the path (LSM3D_PHI_UPWIND_GRAD_F,
LSM3D_AVERAGE_GRAD_PHI, LSM3D_COMPUTE_UNIT_NORMAL,
LSM3D_COMPUTE_SIGNED_UNIT_NORMAL, LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS) are
checked bit for bit against the reference build and its golden vectors in
tests/test_oracle.py.  Here: closed forms of the normals, and the composed
extension (oracle/steps.py:FieldExtensionOracle) against the reference
build's composition and against what the equation fixes -- a planar front
extends the values next to the interface unchanged along the normal, the
cells that straddle the interface do not change (to rounding).
"""
import numpy as np
import pytest

from lsmlib_amd.pylsmlib import BC_LINEAR, Grid
from oracle.oracle import oracle_lib, ref_lib
from oracle.steps import FieldExtensionOracle, fill_ghosts


def plane_case(n=(20, 9, 8), c=0.4376):
    h = 1.0 / n[0]
    grid = Grid(n, (0.0, 0.0, 0.0), (1.0, n[1] * h, n[2] * h), 3, spacing=(h, h, h))
    X, Y, Z = (a.numpy() for a in grid.centers())
    phi = X - c
    S = np.sin(3 * Y) + Z * Z + 0.7 * X + 0.2 * np.cos(5 * X)
    return grid, phi, S


def test_signed_normal_of_a_plane():
    ora = oracle_lib("double")
    grid, phi, _ = plane_case()
    fe = FieldExtensionOracle(ora, grid, (BC_LINEAR,) * 6)
    nx, ny, nz = (grid.interior(a) for a in fe.normals(phi))
    p = grid.interior(phi)
    h = grid.dx[0]
    np.testing.assert_allclose(nx, p / np.sqrt(p * p + h * h), rtol=1e-14)
    assert np.all(ny == 0) and np.all(nz == 0)


def test_unit_normal_and_average_gradient_closed_forms():
    ora = oracle_lib("double")
    gb = (-1, 6, -1, 5, -1, 4)
    fb = (0, 5, 0, 4, 0, 3)
    shape = (6, 7, 8)
    rng = np.random.default_rng(1)
    g = [rng.standard_normal(shape) for _ in range(3)]
    g[0][0] = g[1][0] = g[2][0] = 0.0          # zero gradient plane -> (1, 0, 0)
    n = [np.zeros(shape) for _ in range(3)]
    ora.lsm3dcomputeunitnormal_(*n, gb, *g, gb, fb)
    norm = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
    sl = (slice(1, 5), slice(1, 6), slice(1, 7))
    for d in range(3):
        want = np.where(norm >= 1e-11, g[d] / np.where(norm > 0, norm, 1), 1.0 if d == 0 else 0.0)
        np.testing.assert_allclose(n[d][sl], want[sl], rtol=1e-15)
    avg = [np.zeros(shape) for _ in range(3)]
    ora.lsm3daveragegradphi_(*avg, gb, *g, gb, *n, gb, fb)
    for d in range(3):
        assert np.array_equal(avg[d][sl], (0.5 * (g[d] + n[d]))[sl])


def test_phi_upwind_gradient_picks_the_side_towards_the_interface():
    """phi > 0 takes the derivative from the side where phi decreases, phi <
    0 from the side where it increases; phi == 0 leaves F unchanged."""
    ora = oracle_lib("double")
    gb, fb = (-1, 4, -1, 2, -1, 2), (0, 3, 0, 1, 0, 1)
    shape = (4, 4, 6)
    phi = np.zeros(shape)
    phi[...] = np.arange(-1, 5)[None, None, :] - 1.5   # increasing in x
    phi[1, 1, 3] = 0.0
    Fp = [np.full(shape, 10.0), np.full(shape, 20.0), np.full(shape, 30.0)]
    Fm = [np.full(shape, -10.0), np.full(shape, -20.0), np.full(shape, -30.0)]
    F = [np.full(shape, 7.0) for _ in range(3)]
    ora.lsm3dphiupwindgradf_(*F, gb, *Fp, gb, *Fm, gb, phi, gb, fb)
    x = F[0][1:3, 1:3, 1:5]
    p = phi[1:3, 1:3, 1:5]
    assert np.all(x[p > 0] == -10.0)          # minus side is smaller
    assert np.all(x[p < 0] == 10.0)           # plus side is larger
    assert F[0][1, 1, 3] == 7.0 and F[1][1, 1, 3] == 7.0


def test_extension_of_a_plane_converges_to_the_interface_values():
    ora = oracle_lib("double")
    grid, phi, S = plane_case()
    fe = FieldExtensionOracle(ora, grid, (BC_LINEAR,) * 6)
    S0 = S.copy()
    fill_ghosts(ora, S, grid, fe.bc)
    n, delta = fe.extend(S, phi, max_iterations=400, stop_tol=1e-12)
    assert delta <= 1e-12 and n < 400
    Si, S0i, p = grid.interior(S), grid.interior(S0), grid.interior(phi)
    ineg = np.max(np.nonzero(p[0, 0, :] < 0)[0])     # last column with phi < 0
    # the two columns straddling the interface are frozen (zero right-hand
    # side; the RK3 weights 3/4 + 1/4 and 1/3 + 2/3 round, so up to ulps)
    np.testing.assert_allclose(Si[:, :, ineg:ineg + 2], S0i[:, :, ineg:ineg + 2], rtol=0, atol=1e-13)
    # every other column carries the value of the frozen column on its side
    np.testing.assert_allclose(Si[:, :, :ineg], np.broadcast_to(S0i[:, :, ineg:ineg + 1], Si[:, :, :ineg].shape),
                               atol=1e-10)
    np.testing.assert_allclose(Si[:, :, ineg + 2:],
                               np.broadcast_to(S0i[:, :, ineg + 1:ineg + 2], Si[:, :, ineg + 2:].shape), atol=1e-10)


@pytest.mark.parametrize("real", ["double", "float"])
def test_extension_oracle_matches_reference_build(real):
    ref = ref_lib(real)
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    n = 14
    grid = Grid((n, n - 2, n - 3), (-1.0,) * 3, (1.0, 1.0 - 4.0 / n, 1.0 - 6.0 / n), 3,
                spacing=(2.0 / n,) * 3)
    X, Y, Z = (a.numpy() for a in grid.centers())
    dt = np.float64 if real == "double" else np.float32
    phi = (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.5).astype(dt)
    S = (np.sin(2 * X) * Y + Z).astype(dt)
    outs = []
    for lib in (oracle_lib(real), ref):
        s = S.copy()
        fe = FieldExtensionOracle(lib, grid, (BC_LINEAR,) * 6)
        fe.extend(s, phi, max_iterations=4)
        outs.append(s)
    assert np.array_equal(outs[0], outs[1])
