"""Torch-facing solvers over the gfx950 library (device-resident fields).

Fields are torch tensors of shape (nk, nj, ni) over a ghost box (i fastest,
the reference's column-major layout), on the GPU.  The drivers here sequence
the fused Part-3 entry points of include/lsm3d_amd.h the way LSMLIB's
LevelSetFunctionIntegrator and ReinitializationAlgorithm sequence the toolbox
routines (LevelSetFunctionIntegrator.cc:1319-1510,
ReinitializationAlgorithm.cc:177-362): ghost fill, one fused stage, repeat.
"""
import ctypes
from dataclasses import dataclass

import torch

from ._native import toolbox

RK1, TVDRK2_STAGE2, TVDRK3_STAGE2, TVDRK3_STAGE3 = 1, 22, 32, 33

# boundary-condition codes (BoundaryConditionModule.h:125-130) plus periodic
BC_NONE, BC_NEUMANN, BC_LINEAR, BC_SIGNED_LINEAR, BC_PERIODIC = 0, 1, 2, 3, 5

_DTYPES = {"double": torch.float64, "float": torch.float32}


def real_of(t):
    if t.dtype == torch.float64:
        return "double"
    if t.dtype == torch.float32:
        return "float"
    raise TypeError(f"unsupported dtype {t.dtype}")


def box_c(b):
    return (ctypes.c_int * 6)(*[int(v) for v in b])


def grow(box, g):
    return (box[0] - g, box[1] + g, box[2] - g, box[3] + g, box[4] - g, box[5] + g)


@dataclass
class Grid:
    """Uniform grid: n interior cells per axis on [lower, upper], `ghost`
    ghost layers, fill box indices 0..n-1 (the serial package's convention,
    lsm_grid.c:788-811).

    align (elements, 0 = off): widen the x ghost region so that every row of
    the ghost box, and the first fill-box cell of every row, start on an
    `align`-element boundary (32 for FP32 and 16 for FP64 put them on
    128-byte cache-line boundaries of a 256-byte aligned allocation), with
    an odd number of such blocks per row, so that consecutive rows start in
    different HBM channels (a plane march over rows of 17 cache lines streams
    4 % faster than over rows of 18 on MI355X).  The kernels still read only
    `ghost` cells beyond the fill box; the extra cells are padding.

    spacing: the grid spacing per axis when given (the serial package's
    createGridSetDx takes dx itself, lsm_grid.c:50-114); otherwise
    (upper - lower) / n."""
    n: tuple
    lower: tuple
    upper: tuple
    ghost: int = 3
    align: int = 0
    spacing: tuple = None

    @property
    def xlo(self):  # ghost cells below the fill box along x
        return max(self.ghost, self.align) if self.align else self.ghost

    @property
    def xhi(self):  # ghost cells above the fill box along x
        if not self.align:
            return self.ghost
        hi = self.ghost + (-(self.n[0] + self.xlo + self.ghost)) % self.align
        if ((self.xlo + self.n[0] + hi) // self.align) % 2 == 0:
            hi += self.align  # an odd number of aligned blocks (cache lines) per row
        return hi

    @property
    def fb(self):
        return (0, self.n[0] - 1, 0, self.n[1] - 1, 0, self.n[2] - 1)

    @property
    def gb(self):
        g = self.ghost
        return (-self.xlo, self.n[0] - 1 + self.xhi, -g, self.n[1] - 1 + g, -g, self.n[2] - 1 + g)

    @property
    def shape(self):
        g = self.ghost
        return (self.n[2] + 2 * g, self.n[1] + 2 * g, self.n[0] + self.xlo + self.xhi)

    @property
    def dx(self):
        if self.spacing is not None:
            return tuple(float(h) for h in self.spacing)
        return tuple((self.upper[d] - self.lower[d]) / self.n[d] for d in range(3))

    @property
    def interior_cells(self):
        return self.n[0] * self.n[1] * self.n[2]

    def centers(self, dtype=torch.float64, device="cpu"):
        """Cell-centre coordinates X, Y, Z over the ghost box, each (nk, nj, ni)."""
        gb = self.gb
        ax = [self.lower[d] + (torch.arange(gb[2 * d], gb[2 * d + 1] + 1, dtype=torch.float64) + 0.5) * self.dx[d]
              for d in range(3)]
        Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        return X.to(dtype=dtype, device=device), Y.to(dtype=dtype, device=device), Z.to(dtype=dtype, device=device)

    def interior(self, a):
        g = self.ghost
        return a[g:a.shape[0] - g, g:a.shape[1] - g, self.xlo:self.xlo + self.n[0]]


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("fused stages need device tensors")
    if not t.is_contiguous():
        raise ValueError("tensors must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class Stepper:
    """Device-resident TVD Runge-Kutta drivers for one grid and precision.

    bc: 6 boundary-condition codes (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi).
    ghosts="auto" (default) picks "images" for all-periodic boundaries and
    "synth" otherwise (synth needs no valid ghost cells on entry).
    ghosts="images": the stages read stored ghost cells and write the face
    ghost layers of their output -- periodic images, or the extrapolated
    layers a face fill stores (edge and corner ghost cells of extrapolated
    faces, which no stage reads, are not maintained) -- so the input needs
    valid ghosts once, before the first stage (fill_ghosts).
    ghosts="synth": each stage synthesises the ghost values it reads from
    the fill box and ghost cells of the outputs are not maintained.
    ghosts="fill": a device ghost fill runs after every stage and the stages
    read stored ghost cells (the reference's fill-then-differentiate order).
    Results are identical in every mode.

    zero_tol: threshold of the upwind switch (|V| < zero_tol: no advection
    term) and of the reinitialization cut-off; None keeps the build's
    LSMLIB_ZERO_TOL (1e-11 double, 1e-5 float).  1e-11 in the float library
    runs the double build's switch in FP32 arithmetic.

    tiling: "auto" applies the autotuned (tile shape, z-chunk) for the stage
    kind, precision and grid size (lsmlib_amd.tuning); a (cfg, kz) pair fixes
    it; None leaves the library's current setting.
    """

    def __init__(self, grid, real="double", bc=(BC_PERIODIC,) * 6, device="cuda", ghosts="auto", zero_tol=None,
                 tiling="auto"):
        if ghosts == "auto":
            ghosts = "images" if all(int(b) == BC_PERIODIC for b in bc) else "synth"
        if ghosts not in ("synth", "fill", "images"):
            raise ValueError(ghosts)
        self.grid = grid
        self.real = real
        self.dtype = _DTYPES[real]
        self.lib = toolbox(real)
        self.bc = tuple(int(b) for b in bc)
        self.device = torch.device(device)
        self.ghosts = ghosts
        self._gb = box_c(grid.gb)
        self._fb = box_c(grid.fb)
        self._bc = box_c(self.bc)
        self.all_periodic = all(b == BC_PERIODIC for b in self.bc)
        self._stage_bc = self._bc if ghosts in ("synth", "images") else None
        # bit 8: bc gives the output's face ghost codes, the stage input's ghosts are stored
        self.periodic_mask = 8 if ghosts == "images" else 0
        self.zero_tol = zero_tol
        if tiling == "auto":
            from .tuning import choose_tiling
            n = max(grid.n)
            self._tiling = {k: choose_tiling(real, n, k) for k in ("advect", "reinit")}
        elif tiling is None:
            self._tiling = None
        else:
            cfg, kz = (int(v) for v in tiling)
            self._tiling = {"advect": (cfg, kz), "reinit": (cfg, kz)}
        self.sync(False)

    def sync(self, flag):
        self.lib.set_sync(flag)
        self.lib.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def empty(self):
        return torch.empty(self.grid.shape, dtype=self.dtype, device=self.device)

    def fill_ghosts(self, phi):
        self.lib.fused("lsm3d_amd_fill_ghosts", _ptr(phi), self._gb, self._fb, self._bc)

    def _prefill(self, phi):
        if self.ghosts == "fill":
            self.fill_ghosts(phi)

    # -- advection ---------------------------------------------------------
    def _configure(self, kind):
        self.lib.set_zero_tol(self.zero_tol)
        if self._tiling is not None:
            self.lib.lib.lsm3d_amd_set_stage_tiling(*self._tiling[kind])

    def advect_stage(self, form, out, ps, pc, vel, dt):
        dx, dy, dz = self.grid.dx
        self._configure("advect")
        if isinstance(vel, (tuple, list)) and len(vel) == 3 and torch.is_tensor(vel[0]):
            self.lib.fused("lsm3d_amd_advect_stage", form, _ptr(out), self._gb, _ptr(ps), self._gb,
                           _ptr(pc), self._gb, _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), self._gb,
                           self._fb, dx, dy, dz, dt, self._stage_bc, self.periodic_mask)
        else:
            u, v, w = (float(c) for c in vel)
            self.lib.fused("lsm3d_amd_advect_const_stage", form, _ptr(out), self._gb, _ptr(ps), self._gb,
                           _ptr(pc), self._gb, u, v, w, self._fb, dx, dy, dz, dt, self._stage_bc,
                           self.periodic_mask)

    def advect_lf_stage(self, form, out, ps, pc, vel, dt, alpha=None):
        """One fused stage with the Lax-Friedrichs advection Hamiltonian on the
        plus/minus HJ-WENO5 derivatives: alpha None = local (|v_a| per cell),
        else the global (alpha_x, alpha_y, alpha_z)."""
        dx, dy, dz = self.grid.dx
        self._configure("advect")
        mode, al = self._alpha(alpha)
        self.lib.fused("lsm3d_amd_advect_lf_stage", form, _ptr(out), self._gb, _ptr(ps), self._gb, _ptr(pc), self._gb,
                       _ptr(vel[0]), _ptr(vel[1]), _ptr(vel[2]), self._gb, self._fb, dx, dy, dz, dt, self._stage_bc,
                       self.periodic_mask, mode, al)

    def _alpha(self, alpha):
        if alpha is None:
            return 0, None
        al = (self.lib._ct * 3)(*[float(a) for a in alpha])
        return 1, al

    def advect_lf_rk3_step(self, phi, s1, s2, vel, dt, alpha=None):
        """phi <- one WENO5 / TVD-RK3 step with the Lax-Friedrichs Hamiltonian."""
        self.advect_lf_stage(RK1, s1, phi, None, vel, dt, alpha)
        self._prefill(s1)
        self.advect_lf_stage(TVDRK3_STAGE2, s2, s1, phi, vel, dt, alpha)
        self._prefill(s2)
        self.advect_lf_stage(TVDRK3_STAGE3, phi, s2, phi, vel, dt, alpha)
        self._prefill(phi)

    def advect_rk3_step(self, phi, s1, s2, vel, dt):
        """phi <- one WENO5 / TVD-RK3 step of phi_t + V.grad(phi) = 0 (in place).
        phi's ghost cells must be valid on entry; they are valid on exit."""
        self.advect_stage(RK1, s1, phi, None, vel, dt)
        self._prefill(s1)
        self.advect_stage(TVDRK3_STAGE2, s2, s1, phi, vel, dt)
        self._prefill(s2)
        self.advect_stage(TVDRK3_STAGE3, phi, s2, phi, vel, dt)
        self._prefill(phi)

    def advect_rk2_step(self, phi, s1, vel, dt):
        self.advect_stage(RK1, s1, phi, None, vel, dt)
        self._prefill(s1)
        self.advect_stage(TVDRK2_STAGE2, phi, s1, phi, vel, dt)
        self._prefill(phi)

    # -- reinitialization --------------------------------------------------
    def reinit_stage(self, form, out, ps, pc, phi0, use_phi0, dt):
        dx, dy, dz = self.grid.dx
        self._configure("reinit")
        self.lib.fused("lsm3d_amd_reinit_stage", form, _ptr(out), self._gb, _ptr(ps), self._gb, _ptr(pc),
                       self._gb, _ptr(phi0) if phi0 is not None else None, self._gb, int(use_phi0),
                       self._fb, dx, dy, dz, dt, self._stage_bc, self.periodic_mask)

    def reinit_rk3_iteration(self, phi, s1, s2, dt, phi0=None, use_phi0=False):
        """One TVD-RK3 pseudo-time iteration of the reinitialization equation."""
        if use_phi0 and phi0 is None:
            raise ValueError("use_phi0 needs phi0")
        self.reinit_stage(RK1, s1, phi, None, phi0, use_phi0, dt)
        self._prefill(s1)
        self.reinit_stage(TVDRK3_STAGE2, s2, s1, phi, phi0, use_phi0, dt)
        self._prefill(s2)
        self.reinit_stage(TVDRK3_STAGE3, phi, s2, phi, phi0, use_phi0, dt)
        self._prefill(phi)


def stable_advection_dt(vel, grid, cfl, real=None):
    """LSM3D_COMPUTE_STABLE_ADVECTION_DT over the fill box (device tensors)."""
    real = real or real_of(vel[0])
    lib = toolbox(real)
    dx, dy, dz = grid.dx
    return lib.lsm3dcomputestableadvectiondt_(vel[0], vel[1], vel[2], grid.gb, grid.fb, dx, dy, dz, cfl)


def volume_phi_less_than_zero(phi, grid, eps, real=None):
    """LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO (the interface mass)."""
    real = real or real_of(phi)
    lib = toolbox(real)
    dx, dy, dz = grid.dx
    return lib.lsm3dvolumeregionphilessthanzero_(phi, grid.gb, grid.fb, dx, dy, dz, eps)
