"""Loaders for the CPU checkers of the level-set update path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the lsmlib_amd product.

  oracle_lib(real)  the C restatement (lsm_oracle.c), symbols oracle_<f77>
  ref_lib(real)     the flang build of the reference Fortran in _ref/, when it
                    has been built (None otherwise)
  build(...)        runs this directory's Makefile
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

from lsmlib_amd.f77 import F77Library  # noqa: E402

REFERENCE_TREE = Path(os.environ.get("LSMLIB_REFERENCE", "/root/reference"))


def oracle_path(real):
    return HERE / f"liboracle_{real}.so"


def ref_path(real):
    return HERE / "_ref" / f"libref_{real}.so"


def build(with_reference=None):
    """Compile the C restatement, and the reference Fortran when the
    reference tree is present (or `with_reference` is True)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)
    if with_reference is None:
        with_reference = (REFERENCE_TREE / "src" / "toolbox").is_dir()
    if with_reference:
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REF={REFERENCE_TREE}"], check=True)


def oracle_lib(real="double"):
    p = oracle_path(real)
    if not p.exists():
        build(with_reference=False)
    return F77Library(p, real=real, prefix="oracle_")


def ref_lib(real="double"):
    p = ref_path(real)
    if not p.exists():
        return None
    return F77Library(p, real=real, prefix="")
