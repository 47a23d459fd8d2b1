"""The C-ABI libraries load and export every symbol include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from lsmlib_amd import SIGNATURES, lib_path

REPO = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:void|int|double|float|long\s+long|char|LSMLIB_REAL|lsm3d_amd_\w+)"
                             r"\s*\*?\s*(\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_every_toolbox_symbol():
    decl = declared_symbols()
    missing = set(SIGNATURES) - decl
    assert not missing, f"not declared in include/: {sorted(missing)}"


@pytest.mark.parametrize("real", ["double", "float"])
def test_library_exports_declared_symbols(real):
    p = lib_path(real)
    assert p.exists(), f"{p} not built (python -m lsmlib_amd.build)"
    lib = ctypes.CDLL(str(p))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, f"{p.name} lacks {missing}"


def test_every_3d_routine_of_the_reference_headers_is_exported():
    """Every LSM3D_* macro the reference headers define maps to a symbol of the
    signature table (and so of the library).  Reads the reference tree where
    it is mounted; skipped elsewhere."""
    from lsmlib_amd.signatures import MACRO_OF
    src = Path("/root/reference/src")
    if not src.is_dir():
        pytest.skip("reference tree not mounted")
    macros = {}
    for h in src.rglob("*.h"):
        for m in re.finditer(r"#define\s+(LSM3D_\w+)\s*\\?\s*\n?\s*(lsm3d\w+_)", h.read_text(errors="replace")):
            macros[m.group(1)] = m.group(2).lower()
    assert len(macros) > 100
    have = {v: k for k, v in MACRO_OF.items()}
    missing = sorted(m for m in macros if m not in have)
    assert not missing, f"reference 3D routines not exported: {missing}"
    wrong = sorted(m for m, sym in macros.items() if have.get(m) != sym)
    assert not wrong, f"macro -> symbol differs from the reference: {wrong}"
