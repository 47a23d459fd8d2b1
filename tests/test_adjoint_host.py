"""Host-side pieces of the reverse-mode path (no GPU): the periodic ghost
fold is the transpose of the oracle's periodic ghost fill, and the Fourier
velocity workload is divergence-free and linear in its coefficients."""
import numpy as np
import torch

from lsmlib_amd.adjoint import fold_periodic_ghosts
from lsmlib_amd.pylsmlib import Grid
from lsmlib_amd.workloads import FourierVelocity
from oracle.steps import periodic_fill_axis


def test_fold_is_transpose_of_periodic_fill():
    rng = np.random.default_rng(0)
    g = 3
    shape = (9 + 2 * g, 7 + 2 * g, 8 + 2 * g)
    x = rng.normal(size=shape)
    y = rng.normal(size=shape)
    fx = x.copy()
    for axis in (2, 1, 0):  # the order the faces are filled in (x, y, z)
        periodic_fill_axis(fx, g, axis)
    fy = fold_periodic_ghosts(torch.tensor(y), g).numpy()
    inner = (slice(g, -g),) * 3
    # <fill(x), y> over the ghost box == <x_interior, fold(y)_interior>
    assert np.isclose((fx * y).sum(), (x[inner] * fy[inner]).sum(), rtol=1e-13)
    ghost = fy.copy()
    ghost[inner] = 0
    assert not ghost.any()


def test_fourier_velocity_modes_are_divergence_free():
    grid = Grid((8, 8, 8), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 3)
    fv = FourierVelocity(grid, n_modes=16, seed=4)
    k = torch.tensor(fv.k, dtype=torch.float64)
    assert torch.all(k.abs().max(1).values >= 1) and torch.all(k.abs().max(1).values <= 4)
    # k . P(k) = 0, P symmetric and idempotent
    assert float(torch.einsum("ma,mab->mb", k, fv.proj).abs().max()) < 1e-14
    assert torch.allclose(fv.proj, fv.proj.transpose(1, 2))
    assert torch.allclose(torch.matmul(fv.proj, fv.proj), fv.proj, atol=1e-14)
    # linear in c
    c1, c2 = torch.randn(16, 3, dtype=torch.float64), torch.randn(16, 3, dtype=torch.float64)
    va, vb, vs = fv.velocity(c1), fv.velocity(c2), fv.velocity(2 * c1 - c2)
    for a, b, s in zip(va, vb, vs):
        assert torch.allclose(s, 2 * a - b, atol=1e-13)
    # mode m alone gives P(k_m) c_m sin(pi k_m.x + theta_m)
    X, Y, Z = grid.centers()
    c = torch.zeros(16, 3, dtype=torch.float64)
    c[5] = torch.tensor([0.3, -0.2, 0.7], dtype=torch.float64)
    v = fv.velocity(c)
    kx, ky, kz = (float(t) for t in fv.k[5])
    s = torch.sin(np.pi * (kx * X + ky * Y + kz * Z) + float(fv.theta[5]))
    pc = fv.proj[5] @ c[5]
    for a in range(3):
        assert torch.allclose(v[a], pc[a] * s, atol=1e-14)
