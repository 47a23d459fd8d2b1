"""The autotuned tiling table and its lookup (host logic; no GPU).

Every entry names a compiled tile shape and a z-chunk depth the tuner
sweeps, and choose_tiling falls back as documented: a size class without an
entry takes the nearest tuned class below it, else the one above, else the
default.
"""
import ctypes
import json

import pytest

import lsmlib_amd.tuning as tuning
from lsmlib_amd import lib_path


def tiling_count():
    return ctypes.CDLL(str(lib_path("float"))).lsm3d_amd_stage_tiling_count()


def test_every_table_entry_is_a_compiled_shape_and_a_swept_depth():
    n = tiling_count()
    assert n == 9
    table = json.loads(tuning.TABLE.read_text())
    assert table
    for key, (cfg, kz) in table.items():
        kind = key.split("/")[0]
        assert kind in ("advect", "reinit", "curvature", "patches"), key
        assert 0 <= cfg < n, key
        choices = tuning.CURVATURE_KZ_CHOICES if kind == "curvature" else tuning.KZ_CHOICES
        assert kz in choices, key


def test_choose_tiling_uses_the_entry_of_the_size_class():
    table = json.loads(tuning.TABLE.read_text())
    assert tuning.choose_tiling("float", 512) == tuple(table["advect/float/512"])
    assert tuning.choose_tiling("float", 300) == tuple(table["advect/float/512"])  # class 512
    assert tuning.choose_tiling("float", 0, "patches") == tuple(table["patches/float"])


def test_choose_tiling_falls_back_below_then_above_then_default(tmp_path, monkeypatch):
    t = {"advect/float/128": [3, 8], "advect/float/512": [2, 24], "reinit/float/256": [6, 8]}
    p = tmp_path / "table.json"
    p.write_text(json.dumps(t))
    monkeypatch.setattr(tuning, "TABLE", p)
    assert tuning.choose_tiling("float", 256) == (3, 8)      # nearest class below
    assert tuning.choose_tiling("float", 2048) == (2, 24)
    assert tuning.choose_tiling("float", 64) == (3, 8)       # none below: the one above
    assert tuning.choose_tiling("float", 128, "reinit") == (6, 8)
    assert tuning.choose_tiling("double", 512) == tuning.DEFAULT
    monkeypatch.setattr(tuning, "TABLE", tmp_path / "absent.json")
    assert tuning.choose_tiling("float", 512) == tuning.DEFAULT


@pytest.mark.parametrize("n,cls", [(1, 32), (32, 32), (33, 64), (512, 512), (513, 1024), (5000, 2048)])
def test_size_classes(n, cls):
    assert tuning._size_class(n) == cls
