"""Prototype compatibility with the reference headers, as a C compiler sees it.

One C translation unit includes every 3D toolbox header of the reference
(src/toolbox/*/lsm_*3d*.h) and this library's include/lsm3d_amd.h.  A C
compiler rejects two declarations of one function whose parameter types
differ ("conflicting types"), so compiling it checks every exported
prototype against the reference's own.  The unit also calls a chain of
routines through the reference's LSM3D_* macros and is linked against the
library, so the macro names resolve to exported symbols.  Nothing is run
(that needs a GPU).

The reference headers include LSMLIB_config.h, which configure generates
from config/LSMLIB_config.h.in; every value in that template sits behind
#ifndef, so the test names the template itself LSMLIB_config.h (a link in a
temporary directory) and passes the values configure.ac:258-281 sets for
each precision on the command line.  Skipped where the reference tree is not
mounted.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

from lsmlib_amd import lib_path

REPO = Path(__file__).resolve().parents[1]
REF = Path(os.environ.get("LSMLIB_REFERENCE", "/root/reference"))
CONFIG = {
    "double": ["-DLSMLIB_DOUBLE_PRECISION", "-DLSMLIB_REAL=double", "-DLSMLIB_ZERO_TOL=1.e-11",
               "-DLSMLIB_REAL_MAX=DBL_MAX", "-DLSMLIB_REAL_MIN=DBL_MIN", "-DLSMLIB_REAL_EPSILON=DBL_EPSILON"],
    "float": ["-DLSMLIB_REAL=float", "-DLSMLIB_ZERO_TOL=1.e-5", "-DLSMLIB_REAL_MAX=FLT_MAX",
              "-DLSMLIB_REAL_MIN=FLT_MIN", "-DLSMLIB_REAL_EPSILON=FLT_EPSILON"],
}

CALLS = r"""
#include <float.h>
#include <stdlib.h>
%(includes)s
#include "lsm3d_amd.h"

/* one WENO5 / TVD-RK3 stage as LevelSetMethodToolbox.cc sequences it */
void stage(LSMLIB_REAL *phi, LSMLIB_REAL *u, LSMLIB_REAL *v, LSMLIB_REAL *w, LSMLIB_REAL *gx,
           LSMLIB_REAL *gy, LSMLIB_REAL *gz, LSMLIB_REAL *D1, LSMLIB_REAL *rhs, LSMLIB_REAL *out,
           const int *lo, const int *hi, const int *flo, const int *fhi, LSMLIB_REAL h, LSMLIB_REAL dt) {
  LSMLIB_REAL cfl = 0.5, step;
  LSM3D_COMPUTE_STABLE_ADVECTION_DT(&step, u, v, w, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2,
                                    flo, fhi, flo + 1, fhi + 1, flo + 2, fhi + 2, &h, &h, &h, &cfl);
  LSM3D_UPWIND_HJ_WENO5(gx, gy, gz, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2, phi, lo, hi, lo + 1, hi + 1,
                        lo + 2, hi + 2, u, v, w, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2, D1, lo, hi, lo + 1,
                        hi + 1, lo + 2, hi + 2, flo, fhi, flo + 1, fhi + 1, flo + 2, fhi + 2, &h, &h, &h);
  LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS(rhs, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2);
  LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS(rhs, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2, gx, gy, gz, lo, hi, lo + 1,
                                      hi + 1, lo + 2, hi + 2, u, v, w, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2,
                                      flo, fhi, flo + 1, fhi + 1, flo + 2, fhi + 2);
  LSM3D_TVD_RK3_STAGE1(out, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2, phi, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2,
                       rhs, lo, hi, lo + 1, hi + 1, lo + 2, hi + 2, flo, fhi, flo + 1, fhi + 1, flo + 2, fhi + 2,
                       &dt);
}

int main(void) { return 0; }
"""


@pytest.mark.parametrize("real", ["double", "float"])
def test_reference_headers_compile_and_link_against_the_library(real, tmp_path):
    src = REF / "src" / "toolbox"
    if not src.is_dir():
        pytest.skip("reference tree not mounted")
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = lib_path(real)
    assert lib.exists(), f"{lib} not built"
    headers = sorted(src.glob("*/lsm_*3d*.h"))
    assert len(headers) >= 15
    (tmp_path / "LSMLIB_config.h").symlink_to(REF / "config" / "LSMLIB_config.h.in")
    incs = sorted({str(h.parent) for h in headers})
    text = CALLS % {"includes": "\n".join(f'#include "{h.name}"' for h in headers)}
    c = tmp_path / "use_lsm3d.c"
    c.write_text(text)
    cmd = [cc, "-std=c99", "-Wall", "-Werror", "-Wno-unused-variable", "-I", str(tmp_path), *sum([["-I", d] for d in incs], []),
           "-I", str(REPO / "include"), *CONFIG[real], str(c), "-o", str(tmp_path / "use_lsm3d"),
           "-L", str(lib.parent), f"-llsm3d_amd_{real}", f"-Wl,-rpath,{lib.parent}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
