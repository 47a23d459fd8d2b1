"""Builds the gfx950 shared libraries in-tree.

    python -m lsmlib_amd.build            # both precisions
    python -m lsmlib_amd.build double     # one

Outputs lsmlib_amd/lib/liblsm3d_amd_{double,float}.so; objects go to
lsmlib_amd/_build/<precision>/.  Every source is compiled with hipcc for
gfx950 only, with packed FP32 instructions disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["lsm_runtime.hip", "lsm_toolbox.hip", "lsm_fused.hip", "lsm_batched.hip", "lsm_adjoint.hip",
           "lsm_curvature.hip", "lsm_local.hip", "lsm_local_model.hip", "lsm_probe.hip"]

COMMON_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-munsafe-fp-atomics",
    f"-I{REPO / 'include'}", f"-I{CSRC}",
]


def lib_path(real):
    return LIBDIR / f"liblsm3d_amd_{real}.so"


def _compile(src, real):
    out = BUILD / real / (Path(src).stem + ".o")
    out.parent.mkdir(parents=True, exist_ok=True)
    cmd = [HIPCC, *COMMON_FLAGS, f"-DLSM_REAL={real}", "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src} ({real}):\n{r.stderr}")
    return out


def build(reals=("double", "float"), jobs=None):
    sources = [s for s in SOURCES if (CSRC / s).exists()]
    jobs = jobs or min(8, os.cpu_count() or 4)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {(s, r): ex.submit(_compile, s, r) for r in reals for s in sources}
        objs = {k: f.result() for k, f in futs.items()}
    for r in reals:
        out = lib_path(r)
        tmp = out.with_suffix(".so.tmp")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
               *[str(objs[(s, r)]) for s in sources]]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed ({r}):\n{res.stderr}")
        os.replace(tmp, out)
    return [lib_path(r) for r in reals]


if __name__ == "__main__":
    reals = sys.argv[1:] or ["double", "float"]
    for p in build(reals):
        print(p)
