"""Tile-shape selection for the fused stage kernels.

`choose_tiling(real, n)` returns (cfg, kz) for a grid of n cells per axis from
the table tuning_table.json (written by `autotune`), falling back to a
default.  `autotune(...)` times every compiled tile shape and several z-chunk
depths on the GPU for given grid sizes and records the fastest.
"""
import json
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuning_table.json"
DEFAULT = (0, 32)
KZ_CHOICES = (8, 16, 18, 20, 24, 32, 48, 64, 128)
CURVATURE_KZ_CHOICES = (2, 4, 8, 16, 32, 64)


def _size_class(n):
    for c in (32, 64, 128, 256, 512, 1024):
        if n <= c:
            return c
    return 2048


def load_table():
    if TABLE.exists():
        try:
            return json.loads(TABLE.read_text())
        except ValueError:
            return {}
    return {}


def _key(real, n, kind):
    return f"patches/{real}" if kind == "patches" else f"{kind}/{real}/{_size_class(n)}"


def choose_tiling(real, n, kind="advect"):
    """(cfg, kz) for a stage kind ("advect", "reinit", "curvature" (the
    ENO2 / TVD-RK2 curvature-model stages), or "patches" for the batched
    patch level, where n is ignored).  A size class without an entry
    takes the nearest tuned class below it, else the one above, else the
    default."""
    table = load_table()
    t = table.get(_key(real, n, kind))
    if t is None and kind != "patches":
        classes = sorted(int(k.rsplit("/", 1)[1]) for k in table if k.startswith(f"{kind}/{real}/"))
        below = [c for c in classes if c <= _size_class(n)]
        pick = below[-1] if below else (classes[0] if classes else None)
        t = table.get(f"{kind}/{real}/{pick}") if pick is not None else None
    return tuple(t) if t else DEFAULT


def autotune(reals=("float", "double"), sizes=(128, 256, 512), kind="advect", reps=5, verbose=True, path=None):
    """Times one whole TVD-RK3 step (advect: three stage launches, the
    stages reading phi^n as well as the stage field) or one reinitialization
    stage for every (tile shape, kz) with HIP events (median of `reps`) and
    stores the best in the table at `path`."""
    import torch
    from . import toolbox
    from .pylsmlib import RK1, Stepper
    from .workloads import line_elems, sphere_union_taylor_green

    path = Path(path) if path else TABLE
    table = json.loads(path.read_text()) if path.exists() else {}
    dev = torch.device("cuda:0")
    if kind == "curvature":
        return _autotune_curvature(reals, sizes, reps, verbose, path, table)
    for real in reals:
        lib = toolbox(real)
        dtype = torch.float32 if real == "float" else torch.float64
        ncfg = lib.lib.lsm3d_amd_stage_tiling_count()
        for n in sizes:
            grid, phi, vel, bc = sphere_union_taylor_green(n, n_spheres=8, dtype=dtype, device=dev,
                                                           align=line_elems(dtype))
            st = Stepper(grid, real, bc, tiling=None)  # the sweep sets the tiling
            out = st.empty()
            s2 = st.empty()
            best = None
            for cfg in range(ncfg):
                for kz in KZ_CHOICES:
                    lib.lib.lsm3d_amd_set_stage_tiling(cfg, kz)
                    if kind == "advect":
                        run = lambda: st.advect_rk3_step(phi, out, s2, vel, 1e-3)  # noqa: E731
                    else:
                        run = lambda: st.reinit_stage(RK1, out, phi, None, None, False, 1e-3)  # noqa: E731
                    run()
                    times = []
                    for _ in range(reps):
                        e0 = torch.cuda.Event(enable_timing=True)
                        e1 = torch.cuda.Event(enable_timing=True)
                        e0.record()
                        run()
                        e1.record()
                        e1.synchronize()
                        times.append(e0.elapsed_time(e1))
                    ms = sorted(times)[len(times) // 2]
                    if verbose:
                        print(f"  {kind} {real} n={n} cfg={cfg} kz={kz}: {ms:.3f} ms", flush=True)
                    if best is None or ms < best[0]:
                        best = (ms, cfg, kz)
            table[f"{kind}/{real}/{_size_class(n)}"] = [best[1], best[2]]
            if verbose:
                unit = "ms/step" if kind == "advect" else "ms/stage"
                print(f"autotune {kind} {real} n={n}: cfg={best[1]} kz={best[2]} {best[0]:.3f} {unit}",
                      flush=True)
            del phi, vel, out, s2
            torch.cuda.empty_cache()
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return table


def _autotune_curvature(reals, sizes, reps, verbose, path, table):
    """The curvature-model stage (lsmlib_amd.curvature) on the curvature
    example's pore throat at dx = 1.5 / n (n x n x 2n/3 cells)."""
    import torch
    from . import toolbox
    from .curvature import CurvatureFlow
    from .pylsmlib import TVDRK2_STAGE2
    from .workloads import pore_throat

    dev = torch.device("cuda:0")
    for real in reals:
        lib = toolbox(real)
        dtype = torch.float32 if real == "float" else torch.float64
        ncfg = lib.lib.lsm3d_amd_stage_tiling_count()
        for n in sizes:
            grid, phi, mask = pore_throat(1.5 / n, dtype, dev)
            m = CurvatureFlow(grid, real, mask, device=dev, tiling=None)
            m.fill_ghosts(phi, masked=False)
            m.curvature_stage(1, m.s1, phi, None, 1e-4)
            m.fill_ghosts(m.s1)
            out = m.empty()
            best = None
            for cfg in range(ncfg):
                for kz in CURVATURE_KZ_CHOICES:
                    lib.lib.lsm3d_amd_set_stage_tiling(cfg, kz)
                    run = lambda: m.curvature_stage(TVDRK2_STAGE2, out, m.s1, phi, 1e-4)  # noqa: E731
                    run()
                    times = []
                    for _ in range(reps):
                        e0 = torch.cuda.Event(enable_timing=True)
                        e1 = torch.cuda.Event(enable_timing=True)
                        e0.record()
                        run()
                        e1.record()
                        e1.synchronize()
                        times.append(e0.elapsed_time(e1))
                    ms = sorted(times)[len(times) // 2]
                    if verbose:
                        print(f"  curvature {real} n={max(grid.n)} cfg={cfg} kz={kz}: {ms:.4f} ms", flush=True)
                    if best is None or ms < best[0]:
                        best = (ms, cfg, kz)
            table[f"curvature/{real}/{_size_class(max(grid.n))}"] = [best[1], best[2]]
            if verbose:
                print(f"autotune curvature {real} n={max(grid.n)}: cfg={best[1]} kz={best[2]} {best[0]:.4f} ms/stage",
                      flush=True)
            del phi, mask, out, m
            torch.cuda.empty_cache()
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return table


def autotune_patches(reals=("float", "double"), reps=3, verbose=True, path=None, desc=None):
    """Times one captured TVD-RK3 step of the batched patch level (512 uneven
    patches by default) for every (tile shape, kz) and stores the best."""
    import torch
    from . import toolbox
    from .patches import PatchLevel, random_patch_level
    from .workloads import line_elems

    path = Path(path) if path else TABLE
    table = json.loads(path.read_text()) if path.exists() else {}
    desc = desc or random_patch_level()
    for real in reals:
        dtype = torch.float32 if real == "float" else torch.float64
        ncfg = toolbox(real).lib.lsm3d_amd_stage_tiling_count()
        best = None
        for cfg in range(ncfg):
            for kz in KZ_CHOICES:
                lvl = PatchLevel(desc, real, tiling=(cfg, kz), align=line_elems(dtype))
                phi = lvl.initial_phi()
                s1, s2 = lvl.empty_pool(), lvl.empty_pool()
                lvl.capture_rk3(1, phi, s1, s2, lvl.stable_dt(0.5))
                lvl.replay()
                times = []
                for _ in range(reps):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    lvl.replay()
                    e1.record()
                    e1.synchronize()
                    times.append(e0.elapsed_time(e1))
                ms = sorted(times)[len(times) // 2]
                if verbose:
                    print(f"  patches {real} cfg={cfg} kz={kz}: {ms:.3f} ms/step", flush=True)
                if best is None or ms < best[0]:
                    best = (ms, cfg, kz)
                del lvl, phi, s1, s2
        table[_key(real, 0, "patches")] = [best[1], best[2]]
        if verbose:
            print(f"autotune patches {real}: cfg={best[1]} kz={best[2]} {best[0]:.3f} ms/step", flush=True)
        torch.cuda.empty_cache()
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return table


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time every stage tiling and record the fastest")
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--sizes", default="128,256,512")
    ap.add_argument("--only", default=None, help="advect, reinit, curvature or patches")
    a = ap.parse_args()
    sizes = tuple(int(x) for x in a.sizes.split(","))
    if a.only in (None, "advect"):
        autotune(sizes=sizes, path=a.out)
    if a.only in (None, "reinit"):
        autotune(kind="reinit", sizes=(256,), path=a.out)
    if a.only in (None, "curvature"):
        autotune(kind="curvature", sizes=(38, 75, 256, 512), path=a.out)
    if a.only in (None, "patches"):
        autotune_patches(path=a.out)
