"""Tile-shape selection for the fused stage kernels.

`choose_tiling(real, n)` returns (cfg, kz) for a grid of n cells per axis from
the table tuning_table.json (written by `autotune`), falling back to a
default.  `autotune(...)` times every compiled tile shape and several z-chunk
depths on the GPU for given grid sizes and records the fastest.
"""
import json
import time
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuning_table.json"
DEFAULT = (0, 32)
KZ_CHOICES = (8, 16, 32, 64, 128)


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


def choose_tiling(real, n, kind="advect"):
    t = load_table().get(f"{kind}/{real}/{_size_class(n)}")
    return tuple(t) if t else DEFAULT


def autotune(reals=("float", "double"), sizes=(128, 256, 512), kind="advect", reps=3, verbose=True):
    """Times one stage of `kind` for every (tile shape, kz) and stores the best."""
    import torch
    from . import toolbox
    from .pylsmlib import RK1, Stepper
    from .workloads import sphere_union_taylor_green

    table = load_table()
    dev = torch.device("cuda:0")
    for real in reals:
        lib = toolbox(real)
        dtype = torch.float32 if real == "float" else torch.float64
        ncfg = lib.lib.lsm3d_amd_stage_tiling_count()
        for n in sizes:
            grid, phi, vel, bc = sphere_union_taylor_green(n, n_spheres=8, dtype=dtype, device=dev)
            st = Stepper(grid, real, bc)
            out = st.empty()
            best = None
            for cfg in range(ncfg):
                for kz in KZ_CHOICES:
                    lib.lib.lsm3d_amd_set_stage_tiling(cfg, kz)
                    if kind == "advect":
                        run = lambda: st.advect_stage(RK1, out, phi, None, vel, 1e-3)  # noqa: E731
                    else:
                        run = lambda: st.reinit_stage(RK1, out, phi, None, None, False, 1e-3)  # noqa: E731
                    run()
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        run()
                    torch.cuda.synchronize()
                    ms = (time.perf_counter() - t0) * 1e3 / reps
                    if best is None or ms < best[0]:
                        best = (ms, cfg, kz)
            table[f"{kind}/{real}/{_size_class(n)}"] = [best[1], best[2]]
            if verbose:
                print(f"autotune {kind} {real} n={n}: cfg={best[1]} kz={best[2]} {best[0]:.3f} ms/stage",
                      flush=True)
            del phi, vel, out
            torch.cuda.empty_cache()
    TABLE.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return table


if __name__ == "__main__":
    autotune()
    autotune(kind="reinit", sizes=(256,))
