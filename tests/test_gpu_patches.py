"""GPU parity of the batched patch-level advance.

The oracle advances each patch on its own with the reference's per-patch
sequence (signed-linear ghost fill, upwind HJ-WENO5, advection RHS, TVD-RK3
stage; LevelSetMethodToolbox.cc:2446-2592); the GPU advances all patches with
one launch per stage, and again through a captured hipGraph of several
steps.  FP64: 1e-12 relative to max|phi| per patch; graph replays equal the
per-stage launches bit for bit on every replay.
"""
import numpy as np
import pytest
import torch

from lsmlib_amd.patches import PatchLevel, random_patch_level
from lsmlib_amd.pylsmlib import BC_SIGNED_LINEAR, Grid
from lsmlib_amd.workloads import line_elems
from oracle.oracle import oracle_lib
from oracle.steps import OracleStepper, fill_ghosts

pytestmark = pytest.mark.gpu


def small_level(real="double", npatch=7, align=0):
    desc = random_patch_level(npatch=npatch, lo=6, hi=14, seed=17, dx=1.0 / 32)
    return PatchLevel(desc, real, tiling=(0, 8), align=align)


def oracle_advance(level, phi_pool, dt, steps):
    ora = oracle_lib("double")
    out = []
    h = level.dx
    for q in range(level.npatch):
        nx, ny, nz = (int(v) for v in level.desc["ext"][q])
        grid = Grid((nx, ny, nz), (0.0, 0.0, 0.0), (nx * h, ny * h, nz * h), 3)
        xlo = int(level.xlo[q])
        ph = level.patch_view(phi_pool, q)[:, :, xlo - 3:xlo + nx + 3].to(torch.float64).cpu().numpy().copy()
        vel = tuple(np.full(grid.shape, float(level._vel[q, d])) for d in range(3))
        bc = (BC_SIGNED_LINEAR,) * 6
        st = OracleStepper(ora, grid, bc)
        fill_ghosts(ora, ph, grid, bc)
        for _ in range(steps):
            st.advect_rk3_step(ph, vel, dt)
        out.append(grid.interior(ph))
    return out


def gpu_interiors(level, pool):
    return [level.interior_view(pool, q).to(torch.float64).cpu().numpy() for q in range(level.npatch)]


@pytest.mark.parametrize("align", [0, 16])
def test_patch_level_matches_oracle_fp64(align):
    lvl = small_level("double", align=align)
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.5)
    phi, s1, s2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    for _ in range(3):
        lvl.rk3_step(phi, s1, s2, dt)
    torch.cuda.synchronize()
    got = gpu_interiors(lvl, phi)
    want = oracle_advance(lvl, phi0, dt, 3)
    for q, (a, b) in enumerate(zip(got, want)):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), f"patch {q}"


@pytest.mark.parametrize("real,tiling", [("double", "auto"), ("float", "auto"), ("float", (7, 32))])
def test_patch_level_bench_distribution_matches_oracle(real, tiling):
    """16 patches drawn as the 512-patch bench level draws them (extents
    16..96 per axis, seed 3, dx 1/128): tiles away from every patch face
    take the stored-ghost march, so these are compared with the oracle too.
    FP64 1e-12 of max|phi| per patch; FP32 against the FP64 oracle 1e-5."""
    desc = random_patch_level(npatch=16, lo=16, hi=96, seed=3, dx=1.0 / 128)
    dtype = torch.float64 if real == "double" else torch.float32
    lvl = PatchLevel(desc, real, tiling=tiling, align=line_elems(dtype))
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.5)
    phi, s1, s2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    for _ in range(3):
        lvl.rk3_step(phi, s1, s2, dt)
    torch.cuda.synchronize()
    want = oracle_advance(lvl, phi0, dt, 3)
    tol = 1e-12 if real == "double" else 1e-5
    for q, (a, b) in enumerate(zip(gpu_interiors(lvl, phi), want)):
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), f"patch {q} ext {desc['ext'][q]}"


def test_patch_level_fp32_vs_fp64_oracle():
    lvl = small_level("float")
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.5)
    phi, s1, s2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    for _ in range(3):
        lvl.rk3_step(phi, s1, s2, dt)
    torch.cuda.synchronize()
    want = oracle_advance(lvl, phi0, dt, 3)
    for a, b in zip(gpu_interiors(lvl, phi), want):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


@pytest.mark.parametrize("real", ["double", "float"])
def test_graph_replays_match_eager_stages(real):
    lvl = small_level(real, npatch=9)
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.4)
    # eager: 2 steps per replay, 3 replays
    e_phi, e1, e2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    g_phi, g1, g2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    lvl.capture_rk3(2, g_phi, g1, g2, dt)
    for rep in range(3):
        for _ in range(2):
            lvl.rk3_step(e_phi, e1, e2, dt)
        lvl.replay()
        torch.cuda.synchronize()
        _ = (e_phi * 2).sum().item()  # unrelated eager work between replays
        assert torch.equal(g_phi, e_phi), f"replay {rep} differs from eager stages"


def test_tiling_does_not_change_results():
    desc = random_patch_level(npatch=5, lo=6, hi=20, seed=5, dx=1.0 / 40)
    ref = None
    from lsmlib_amd import toolbox
    for cfg in range(toolbox("double").lib.lsm3d_amd_stage_tiling_count()):
        for kz in (3, 16):
            lvl = PatchLevel(desc, "double", tiling=(cfg, kz))
            phi = lvl.initial_phi()
            s1, s2 = lvl.empty_pool(), lvl.empty_pool()
            lvl.rk3_step(phi, s1, s2, 0.003)
            torch.cuda.synchronize()
            out = torch.cat([torch.from_numpy(a.ravel()) for a in gpu_interiors(lvl, phi)])
            if ref is None:
                ref = out
            else:
                assert torch.equal(out, ref), f"tiling {cfg}/{kz}"


@pytest.mark.parametrize("real", ["double", "float"])
def test_fill_ghosts_writes_extrapolated_faces(real):
    """lsm3d_amd_patch_plan_fill_ghosts: every patch's face ghost layers hold
    the signed-linear extrapolation the oracle's face-by-face fill produces
    from the fill box (edges and corners are not written and not compared)."""
    lvl = small_level(real, npatch=6, align=32 if real == "float" else 16)
    phi = lvl.initial_phi()
    lvl.fill_ghosts(phi)
    torch.cuda.synchronize()
    ora = oracle_lib(real)
    for q in range(lvl.npatch):
        nx, ny, nz = (int(v) for v in lvl.desc["ext"][q])
        grid = Grid((nx, ny, nz), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3)
        xlo = int(lvl.xlo[q])
        got = lvl.patch_view(phi, q)[:, :, xlo - 3:xlo + nx + 3].cpu().numpy()
        want = got.copy()
        fill_ghosts(ora, want, grid, (BC_SIGNED_LINEAR,) * 6)
        g = 3
        faces = [(slice(g, -g), slice(g, -g), slice(0, g)), (slice(g, -g), slice(g, -g), slice(-g, None)),
                 (slice(g, -g), slice(0, g), slice(g, -g)), (slice(g, -g), slice(-g, None), slice(g, -g)),
                 (slice(0, g), slice(g, -g), slice(g, -g)), (slice(-g, None), slice(g, -g), slice(g, -g))]
        for f, sl in enumerate(faces):
            assert np.array_equal(got[sl], want[sl]), f"patch {q} face {f}"


def test_stored_face_ghosts_equal_synthesised():
    """Stages with code-0 faces reading ghosts stored by fill_ghosts give the
    same step as stages synthesising the signed-linear ghosts."""
    desc = random_patch_level(npatch=6, lo=6, hi=14, seed=23, dx=1.0 / 32)
    outs = []
    for bc in ((BC_SIGNED_LINEAR,) * 6, (0,) * 6):
        lvl = PatchLevel(desc, "double", tiling=(2, 8), bc=bc)
        phi = lvl.initial_phi()
        s1, s2 = lvl.empty_pool(), lvl.empty_pool()
        fill = PatchLevel(desc, "double", tiling=(2, 8))
        dt = lvl.stable_dt(0.4)
        for _ in range(2):
            if bc[0] == 0:
                fill.fill_ghosts(phi)
            lvl.stage(1, s1, phi, None, dt)
            if bc[0] == 0:
                fill.fill_ghosts(s1)
            lvl.stage(32, s2, s1, phi, dt)
            if bc[0] == 0:
                fill.fill_ghosts(s2)
            lvl.stage(33, phi, s2, phi, dt)
        torch.cuda.synchronize()
        outs.append(torch.cat([torch.from_numpy(a.ravel()) for a in gpu_interiors(lvl, phi)]))
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("real", ["double", "float"])
def test_full_512_patch_level_sampled_patches(real):
    """The benchmarked level itself (512 patches, extents 16..96, seed 3,
    dx 1/128, tuned tiling, one persistent launch per stage through a captured
    hipGraph of 2 steps): 8 patches drawn with a fixed seed are advanced by the
    oracle from the same initial pool.  A patch's stages read only its own
    pool slice and synthesised face ghosts, so each is an independent oracle
    case.  FP64 1e-12 of max|phi| per patch; FP32 against the FP64 oracle 1e-5."""
    desc = random_patch_level(npatch=512, lo=16, hi=96, seed=3, dx=1.0 / 128)
    dtype = torch.float64 if real == "double" else torch.float32
    lvl = PatchLevel(desc, real, align=line_elems(dtype))
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.5)
    phi, s1, s2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    lvl.capture_rk3(2, phi, s1, s2, dt)
    lvl.replay()
    torch.cuda.synchronize()
    picks = np.random.default_rng(11).choice(lvl.npatch, size=8, replace=False)
    want = oracle_advance_patches(lvl, phi0, dt, 2, picks)
    tol = 1e-12 if real == "double" else 1e-5
    for q, b in zip(picks, want):
        a = lvl.interior_view(phi, int(q)).to(torch.float64).cpu().numpy()
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), f"patch {q} ext {desc['ext'][q]}"


def oracle_advance_patches(level, phi_pool, dt, steps, picks):
    """oracle_advance for the patches listed in `picks` only."""
    ora = oracle_lib("double")
    out = []
    h = level.dx
    for q in (int(v) for v in picks):
        nx, ny, nz = (int(v) for v in level.desc["ext"][q])
        grid = Grid((nx, ny, nz), (0.0, 0.0, 0.0), (nx * h, ny * h, nz * h), 3)
        xlo = int(level.xlo[q])
        ph = level.patch_view(phi_pool, q)[:, :, xlo - 3:xlo + nx + 3].to(torch.float64).cpu().numpy().copy()
        vel = tuple(np.full(grid.shape, float(level._vel[q, d])) for d in range(3))
        bc = (BC_SIGNED_LINEAR,) * 6
        st = OracleStepper(ora, grid, bc)
        fill_ghosts(ora, ph, grid, bc)
        for _ in range(steps):
            st.advect_rk3_step(ph, vel, dt)
        out.append(grid.interior(ph))
    return out


def test_captured_stage_pools_stay_alive():
    """The captured graph writes the stage pools it was captured with; the
    level keeps them alive, so pools the caller does not hold are not freed
    and handed to later allocations (which the replay would overwrite)."""
    lvl = small_level("float", npatch=9)
    phi0 = lvl.initial_phi()
    dt = lvl.stable_dt(0.4)
    g_phi = phi0.clone()
    lvl.capture_rk3(2, g_phi, lvl.empty_pool(), lvl.empty_pool(), dt)
    later = [torch.full_like(g_phi, float("nan")) for _ in range(4)]
    lvl.replay()
    torch.cuda.synchronize()
    e_phi, e1, e2 = phi0.clone(), lvl.empty_pool(), lvl.empty_pool()
    for _ in range(2):
        lvl.rk3_step(e_phi, e1, e2, dt)
    torch.cuda.synchronize()
    assert torch.equal(g_phi, e_phi)
    assert all(bool(torch.isnan(t).all()) for t in later)
