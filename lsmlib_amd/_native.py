"""Loads the in-tree gfx950 libraries.

There is no CPU fallback: if the shared library is missing, or no GPU is
visible when a compute entry point is called, the call raises.
"""
import ctypes
from pathlib import Path

try:  # load torch's HIP runtime first so torch and this library share one
    import torch  # noqa: F401  (same SONAME libamdhip64.so.7: first loaded wins)
except ImportError:  # pragma: no cover - torch is optional for the C ABI
    torch = None

from .f77 import F77Library

PKG = Path(__file__).resolve().parent
LIBDIR = PKG / "lib"

_LIBS = {}


class LsmError(RuntimeError):
    """An entry point reported invalid arguments or a HIP failure."""


def lib_path(real):
    return LIBDIR / f"liblsm3d_amd_{real}.so"


class LocalCurvatureModel(ctypes.Structure):
    """LsmLocalCurvatureModel of include/lsm3d_amd.h (the narrow-band
    curvature example's arrays and state)."""
    ARRAYS = ("phi_x_plus", "phi_y_plus", "phi_z_plus", "phi_x_minus", "phi_y_minus", "phi_z_minus", "D1", "D2",
              "phi_x", "phi_y", "phi_z", "phi_xx", "phi_xy", "phi_xz", "phi_yy", "phi_yz", "phi_zz", "lse_rhs",
              "phi_stage1", "phi_next", "phi0")
    _fields_ = ([(n, ctypes.c_void_p) for n in ARRAYS] +
                [("mask", ctypes.c_void_p), ("narrow_band", ctypes.c_void_p), ("index_x", ctypes.c_void_p),
                 ("index_y", ctypes.c_void_p), ("index_z", ctypes.c_void_p), ("index_outer", ctypes.c_void_p),
                 ("gb", ctypes.c_int * 6), ("fb", ctypes.c_int * 6), ("dx", ctypes.c_double * 3),
                 ("a", ctypes.c_double), ("b", ctypes.c_double), ("cfl", ctypes.c_double),
                 ("n_lo", ctypes.c_int * 6), ("n_hi", ctypes.c_int * 6), ("outer_limits", ctypes.c_int * 4)])


class ToolboxLibrary(F77Library):
    """The product library: LSM3D_* entry points plus runtime control.

    Calls run in the library's "record" error mode and raise LsmError with
    the library's message instead of aborting the process.
    """

    def __init__(self, real="double"):
        p = lib_path(real)
        if not p.exists():
            raise LsmError(f"{p} is not built; run `python -m lsmlib_amd.build` "
                           "(there is no CPU fallback)")
        super().__init__(p, real=real, prefix="")
        L = self.lib
        L.lsm_amd_set_error_mode.argtypes = [ctypes.c_int]
        L.lsm_amd_set_stream.argtypes = [ctypes.c_void_p]
        L.lsm_amd_get_stream.restype = ctypes.c_void_p
        L.lsm_amd_set_sync.argtypes = [ctypes.c_int]
        L.lsm_amd_last_error.restype = ctypes.c_char_p
        L.lsm_amd_error_pending.restype = ctypes.c_int
        L.lsm_amd_real_size.restype = ctypes.c_int
        self._bind_fused()
        want = 8 if real == "double" else 4
        if L.lsm_amd_real_size() != want:
            raise LsmError(f"{p} was built for a different LSMLIB_REAL")

    def _bind_fused(self):
        """argtypes of the fused device API (plain values, int[6] boxes)."""
        L, R = self.lib, self._ct
        P, I, B = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        L.lsm3d_amd_advect_stage.argtypes = [I, P, B, P, B, P, B, P, P, P, B, B, R, R, R, R, B, I]
        L.lsm3d_amd_advect_const_stage.argtypes = [I, P, B, P, B, P, B, R, R, R, B, R, R, R, R, B, I]
        L.lsm3d_amd_reinit_stage.argtypes = [I, P, B, P, B, P, B, P, B, I, B, R, R, R, R, B, I]
        L.lsm3d_amd_fill_ghosts.argtypes = [P, B, B, B]
        RP = ctypes.POINTER(R)
        L.lsm3d_amd_compute_stable_advection_dt_noalloc.argtypes = [P, P, P, P, B, B, R, R, R, R]
        L.lsm3d_amd_compute_stable_advection_dt_noalloc.restype = None
        L.lsm3d_amd_advect_lf_stage.argtypes = [I, P, B, P, B, P, B, P, P, P, B, B, R, R, R, R, B, I, I, RP]
        L.lsm3d_amd_add_lf_advection_term.argtypes = [P, B, P, P, P, B, P, P, P, B, P, P, P, B, B, I, RP]
        L.lsm3d_amd_add_lf_normal_vel_term.argtypes = [P, B, P, P, P, B, P, P, P, B, P, B, R, B, I, R]
        L.lsm3d_amd_set_stage_tiling.argtypes = [I, I]
        L.lsm3d_amd_stage_tiling_count.restype = I
        L.lsm3d_amd_stage_tiling_shape.argtypes = [I, B, B]
        LL = ctypes.POINTER(ctypes.c_longlong)
        L.lsm3d_amd_patch_plan_create.argtypes = [I, B, B, LL, ctypes.c_longlong, ctypes.POINTER(R), I, I]
        L.lsm3d_amd_patch_plan_create.restype = P
        L.lsm3d_amd_patch_plan_destroy.argtypes = [P]
        for f in ("cells", "tiles"):
            getattr(L, f"lsm3d_amd_patch_plan_{f}").argtypes = [P]
            getattr(L, f"lsm3d_amd_patch_plan_{f}").restype = ctypes.c_longlong
        L.lsm3d_amd_patch_plan_workgroups.argtypes = [P]
        L.lsm3d_amd_patch_plan_workgroups.restype = I
        L.lsm3d_amd_patch_plan_stage.argtypes = [P, I, P, P, P, R, R, R, R, B]
        L.lsm3d_amd_patch_plan_capture_rk3.argtypes = [P, I, P, P, P, R, R, R, R, B]
        L.lsm3d_amd_patch_plan_replay.argtypes = [P]
        L.lsm3d_amd_patch_plan_fill_ghosts.argtypes = [P, P, B]
        L.lsm_amd_bandwidth_probe.argtypes = [I, P, P, ctypes.c_longlong, I, I]
        L.lsm_amd_bandwidth_probe.restype = ctypes.c_double
        L.lsm_amd_march_probe.argtypes = [P, P, I, I, I, I, I]
        L.lsm_amd_march_probe.restype = ctypes.c_double
        L.lsm_amd_valu_probe.argtypes = [P, I, I, I, I]
        L.lsm_amd_valu_probe.restype = ctypes.c_double
        L.lsm3d_amd_set_zero_tol.argtypes = [ctypes.c_double]
        L.lsm3d_amd_set_zero_tol.restype = None
        L.lsm3d_amd_zero_tol.argtypes = []
        L.lsm3d_amd_zero_tol.restype = ctypes.c_double
        L.lsm3d_amd_advect_adjoint_work_elems.argtypes = [B]
        L.lsm3d_amd_advect_adjoint_work_elems.restype = ctypes.c_longlong
        L.lsm3d_amd_advect_stage_adjoint.argtypes = [I, P, B, P, B, P, B, I, P, B, P, P, P, B, P, P, P, B, P, B,
                                                     R, R, R, R, I]
        L.lsm3d_amd_heaviside_mismatch.argtypes = [P, B, P, B, B, R, P, P, B]
        L.lsm3d_amd_curvature_stage.argtypes = [I, P, B, P, B, P, B, P, B, P, B, B, R, R, R, R, R, R]
        L.lsm3d_amd_reinit_eno2_stage.argtypes = [I, P, B, P, B, P, B, P, B, P, B, I, P, B, B, R, R, R, R]
        L.lsm3d_amd_extrapolate_ghosts_masked.argtypes = [P, B, P, B, P, B, B, B]
        L.lsm3d_amd_const_normal_vel_dt_eno2.argtypes = [ctypes.POINTER(R), P, B, B, R, R, R, R, R]
        L.lsm3d_amd_impose_mask.argtypes = [P, P, P, ctypes.c_longlong]
        L.lsm3d_amd_field_extension_velocity.argtypes = [P, P, P, B, P, P, P, B, P, B, B, R, R, R]
        M = ctypes.POINTER(LocalCurvatureModel)
        D = ctypes.c_double
        L.lsm3d_amd_local_curvature_step.argtypes = [M, P, D, D]
        L.lsm3d_amd_local_curvature_step.restype = D
        L.lsm3d_amd_local_curvature_reinit_trigger.argtypes = [M, P, D, D, D]
        L.lsm3d_amd_local_curvature_reinit_trigger.restype = I
        L.lsm3d_amd_local_curvature_reinit.argtypes = [M, P, D]
        L.lsm3d_amd_local_curvature_reinit.restype = I
        for f in ("lsm3d_amd_advect_stage", "lsm3d_amd_advect_const_stage", "lsm3d_amd_reinit_stage",
                  "lsm3d_amd_advect_lf_stage", "lsm3d_amd_add_lf_advection_term", "lsm3d_amd_add_lf_normal_vel_term",
                  "lsm3d_amd_fill_ghosts", "lsm3d_amd_set_stage_tiling", "lsm3d_amd_patch_plan_destroy",
                  "lsm3d_amd_patch_plan_stage", "lsm3d_amd_patch_plan_capture_rk3",
                  "lsm3d_amd_patch_plan_replay", "lsm3d_amd_patch_plan_fill_ghosts", "lsm3d_amd_advect_stage_adjoint",
                  "lsm3d_amd_heaviside_mismatch", "lsm3d_amd_curvature_stage", "lsm3d_amd_reinit_eno2_stage",
                  "lsm3d_amd_extrapolate_ghosts_masked", "lsm3d_amd_const_normal_vel_dt_eno2",
                  "lsm3d_amd_impose_mask", "lsm3d_amd_field_extension_velocity"):
            getattr(L, f).restype = None

    def fused(self, fname, *args):
        """Calls a fused-API function (Part 3 of include/lsm3d_amd.h)."""
        self._prepare()
        out = getattr(self.lib, fname)(*args)
        self._check(fname)
        return out

    def _prepare(self):
        self.lib.lsm_amd_set_error_mode(1)

    def _check(self, name):
        if self.lib.lsm_amd_error_pending():
            msg = self.lib.lsm_amd_last_error().decode(errors="replace")
            self.lib.lsm_amd_clear_error()
            raise LsmError(msg)

    def call(self, name, *args):
        self._prepare()
        out = super().call(name, *args)
        self._check(name)
        return out

    def set_stream(self, stream_handle):
        self.lib.lsm_amd_set_stream(ctypes.c_void_p(int(stream_handle or 0)))

    def set_sync(self, sync):
        self.lib.lsm_amd_set_sync(int(bool(sync)))

    def set_zero_tol(self, tol=None):
        """Upwind-switch / reinit cut-off of the fused stages on this host
        thread; None restores the build's LSMLIB_ZERO_TOL."""
        self.lib.lsm3d_amd_set_zero_tol(-1.0 if tol is None else float(tol))

    def zero_tol(self):
        return self.lib.lsm3d_amd_zero_tol()


def toolbox(real="double"):
    """The LSM3D_* library for `real` ('double' or 'float'), loaded once."""
    if real not in _LIBS:
        _LIBS[real] = ToolboxLibrary(real)
    return _LIBS[real]
