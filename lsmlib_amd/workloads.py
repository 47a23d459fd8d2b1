"""Synthetic level-set workloads (seeded), built with torch on any device.

  rotating_sphere(n)       sphere r=0.35 centred (0.35,0,0) on [-1,1]^3, rigid
                           rotation u=(-y,x,0), signed-linear ghost fill
                           (advection parity case, CFL 0.5)
  wavy_sphere_reinit(n)    phi0 = (r-0.5)(1+0.4 sin3x sin3y sin3z) on [-1,1]^3
                           (reinitialization case)
  sphere_union_taylor_green(n, n_spheres=64, seed=5)
                           periodic [0,1]^3, phi0 = min of sphere signed
                           distances (radii U[0.02,0.1]), Taylor-Green velocity
                           u=sin2pix cos2piy cos2piz, v=-cos2pix sin2piy cos2piz,
                           w=0 (large advection case)
  align=line_elems(dtype) pads the x ghost region so rows start on 128-byte
  cache-line boundaries (see pylsmlib.Grid)
  FourierVelocity(grid, ...) / adjoint_case(n)
                           divergence-free velocity sum_m P(k_m) c_m
                           sin(pi k_m.x + theta_m), 64 integer wavevectors with
                           1 <= max|k_i| <= 4, c ~ N(0, 0.1^2), on periodic
                           [-1,1]^3; phi0 = sphere r=0.3 at the origin
                           (gradient case)
  pore_throat(dx)          the curvature example's default problem: a planar
                           front x = x_lo + 20 dx in the pore throat between
                           three unit spheres, box [-0.75,0.75]^2 x [0,1],
                           mask = the spheres (positive inside), node
                           coordinates of the serial package
"""
import math

import numpy as np
import torch

from .pylsmlib import BC_PERIODIC, BC_SIGNED_LINEAR, Grid


def line_elems(dtype):
    """Elements per 128-byte cache line (the Grid alignment for `dtype`)."""
    return 128 // torch.empty((), dtype=dtype).element_size()


def rotating_sphere(n=128, dtype=torch.float64, device="cpu", align=0):
    grid = Grid((n, n, n), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 3, align)
    X, Y, Z = grid.centers(torch.float64, device)
    phi = torch.sqrt((X - 0.35) ** 2 + Y ** 2 + Z ** 2) - 0.35
    vel = (-Y, X, torch.zeros_like(X))
    return (grid, phi.to(dtype).contiguous(), tuple(v.to(dtype).contiguous() for v in vel),
            (BC_SIGNED_LINEAR,) * 6)


def wavy_sphere_reinit(n=256, dtype=torch.float64, device="cpu", align=0):
    grid = Grid((n, n, n), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 3, align)
    X, Y, Z = grid.centers(torch.float64, device)
    r = torch.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    phi = (r - 0.5) * (1 + 0.4 * torch.sin(3 * X) * torch.sin(3 * Y) * torch.sin(3 * Z))
    return grid, phi.to(dtype).contiguous(), (BC_SIGNED_LINEAR,) * 6


def taylor_green(grid, dtype=torch.float64, device="cpu"):
    X, Y, Z = grid.centers(torch.float64, device)
    tp = 2 * math.pi
    u = torch.sin(tp * X) * torch.cos(tp * Y) * torch.cos(tp * Z)
    v = -torch.cos(tp * X) * torch.sin(tp * Y) * torch.cos(tp * Z)
    w = torch.zeros_like(u)
    return tuple(a.to(dtype).contiguous() for a in (u, v, w))


def sphere_union(grid, n_spheres=64, seed=5, rmin=0.02, rmax=0.1, dtype=torch.float64, device="cpu",
                 periodic=True):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(n_spheres, 3))
    radii = rng.uniform(rmin, rmax, size=n_spheres)
    X, Y, Z = grid.centers(torch.float64, device)
    phi = torch.full_like(X, float("inf"))
    for c, r in zip(centres, radii):
        d = []
        for A, cc in ((X, c[0]), (Y, c[1]), (Z, c[2])):
            t = A - cc
            if periodic:
                t = t - torch.round(t)  # minimum image on the unit torus
            d.append(t)
        phi = torch.minimum(phi, torch.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) - r)
        del d
    return phi.to(dtype).contiguous()


def abc_flow(grid, dtype=torch.float64, device="cpu"):
    """Arnold-Beltrami-Childress flow with A = B = C = 1 on the unit cube:
    u = sin2piz + cos2piy, v = sin2pix + cos2piz, w = sin2piy + cos2pix
    (divergence-free, every component non-zero almost everywhere)."""
    X, Y, Z = grid.centers(torch.float64, device)
    tp = 2 * math.pi
    u = torch.sin(tp * Z) + torch.cos(tp * Y)
    v = torch.sin(tp * X) + torch.cos(tp * Z)
    w = torch.sin(tp * Y) + torch.cos(tp * X)
    return tuple(a.to(dtype).contiguous() for a in (u, v, w))


def sphere_union_taylor_green(n=512, n_spheres=64, seed=5, dtype=torch.float64, device="cpu", align=0,
                              velocity="taylor_green"):
    """The large advection case; velocity="abc" replaces the Taylor-Green field
    (w = 0) by the fully three-dimensional ABC flow."""
    grid = Grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3, align)
    phi = sphere_union(grid, n_spheres, seed, dtype=dtype, device=device)
    vel = abc_flow(grid, dtype, device) if velocity == "abc" else taylor_green(grid, dtype, device)
    return grid, phi, vel, (BC_PERIODIC,) * 6


class FourierVelocity:
    """V(x) = sum_m P(k_m) c_m sin(pi k_m . x + theta_m), P(k) = I - k k^T/|k|^2
    (divergence-free), as (vx, vy, vz) over the grid's ghost box.  The modes
    (k_m integer, 1 <= max_i |k_mi| <= kmax), phases theta_m ~ U[0, 2 pi)
    and coefficients c ~ N(0, sigma^2) (n_modes x 3) are drawn from `seed`.
    velocity(c) is differentiable in c (a product with the stored basis)."""

    def __init__(self, grid, n_modes=64, kmax=4, sigma=0.1, seed=4, dtype=torch.float64, device="cpu"):
        rng = np.random.default_rng(seed)
        ks = []
        while len(ks) < n_modes:
            k = rng.integers(-kmax, kmax + 1, size=3)
            if np.abs(k).max() >= 1:
                ks.append(k)
        self.k = np.array(ks, dtype=np.int64)
        self.theta = rng.uniform(0.0, 2 * math.pi, size=n_modes)
        self.c0 = torch.tensor(rng.normal(0.0, sigma, size=(n_modes, 3)), dtype=torch.float64)
        self.grid, self.dtype, self.device = grid, dtype, torch.device(device)
        kf = torch.tensor(self.k, dtype=torch.float64)
        proj = torch.eye(3, dtype=torch.float64) - kf[:, :, None] * kf[:, None, :] / (kf ** 2).sum(1)[:, None, None]
        self.proj = proj.to(dtype=dtype, device=self.device)  # (m, 3, 3), symmetric
        X, Y, Z = grid.centers(torch.float64, device)
        basis = torch.empty((n_modes,) + tuple(X.shape), dtype=dtype, device=self.device)
        for m in range(n_modes):
            kx, ky, kz = (float(v) for v in self.k[m])
            basis[m] = torch.sin(math.pi * (kx * X + ky * Y + kz * Z) + float(self.theta[m])).to(dtype)
        self.basis = basis.reshape(n_modes, -1)
        self.shape = tuple(X.shape)

    def velocity(self, c):
        """(vx, vy, vz), each contiguous over the ghost box."""
        q = torch.einsum("mab,mb->ma", self.proj, c.to(self.dtype))  # P(k_m) c_m
        v = torch.matmul(q.t().contiguous(), self.basis)             # (3, cells)
        return tuple(v[a].reshape(self.shape) for a in range(3))


def adjoint_case(n=128, dtype=torch.float64, device="cpu", seed=4):
    """Grid [-1,1]^3 periodic, phi0 = |x| - 0.3 and the Fourier velocity."""
    grid = Grid((n, n, n), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 3)
    X, Y, Z = grid.centers(torch.float64, device)
    phi0 = (torch.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.3).to(dtype).contiguous()
    fv = FourierVelocity(grid, seed=seed, dtype=dtype, device=device)
    return grid, phi0, fv, (BC_PERIODIC,) * 6


def pore_throat(dx=0.04, dtype=torch.float64, device="cpu", align=0):
    """(grid, phi, mask) of the curvature example's default run: the grid of
    createMaskThroatFromSpheres3d (curvature_model_top.c:247-296; ghost width 3
    of the medium-accuracy grid, lsm_grid.c:46,50-114), the mask of
    createIntersectionOfSpheres with inside_flag 1 (lsm_initialization3d.c:115-166),
    the plane of createPlane through (x_lo + 20 dx, 0, 0) with normal (1, 0, 0)
    (lsm_initialization3d.c:18-82) and IMPOSE_MASK of the plane
    (curvature_model_top.c:58-79).  Grid points sit at x_lo_gb + dx * q for
    ghost-box index q (no half-cell offset), as in those routines; values are
    formed in double precision in the routines' operation order."""
    r = 1.0
    x_lo = (-3 * r / 4, -3 * r / 4, 0.0)
    x_hi = (3 * r / 4, 3 * r / 4, r)
    n = tuple(int(np.rint((x_hi[d] - x_lo[d]) / dx)) for d in range(3))
    g = 3
    grid = Grid(n, x_lo, tuple(x_lo[d] + n[d] * dx for d in range(3)), g, align, spacing=(dx, dx, dx))
    gb = grid.gb
    ax = []
    for d in range(3):
        q = np.arange(gb[2 * d], gb[2 * d + 1] + 1, dtype=np.float64) + g  # ghost-box index
        ax.append((x_lo[d] - g * dx) + dx * q)
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    centers = ((0.0, -r, 0.0), (0.0, r, 0.0), (0.0, 0.0, r * math.sqrt(3.0)))
    mask = np.full(X.shape, -np.finfo(np.float32).max, dtype=np.float64)
    for cx, cy, cz in centers:
        sd = -(np.sqrt((X - cx) * (X - cx) + (Y - cy) * (Y - cy) + (Z - cz) * (Z - cz)) - r)
        mask = np.where(sd > mask, sd, mask)
    px = x_lo[0] + 20 * dx
    phi = ((X - px) * 1.0 + (Y - 0.0) * 0.0 + (Z - 0.0) * 0.0) / math.sqrt(1.0)
    phi = np.where(mask > phi, mask, phi)
    to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dtype, device=device)  # noqa: E731
    return grid, to(phi), to(mask)
