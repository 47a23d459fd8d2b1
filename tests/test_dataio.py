"""The serial package's data files (lsmlib_amd/dataio.py).

No reference data files ship with the reference, so the layout is pinned by
building the bytes the serial package's fwrite calls produce
(lsm_data_arrays.c:363-490, lsm_grid.c:604-760) and comparing, and by round
trips (plain, gzip, bzip2) in both precisions.  The grid fields are pinned
against the curvature example's grid (createGridSetDx at dx 0.04 on
[-0.75,0.75]^2 x [0,1] with ghost width 3: 44 x 44 x 31 points).
"""
import gzip
import struct

import numpy as np
import pytest

from lsmlib_amd import dataio


@pytest.mark.parametrize("real,dt", [("double", np.float64), ("float", np.float32)])
@pytest.mark.parametrize("zip", [None, "gzip", "bzip2"])
def test_data_array_round_trip(tmp_path, real, dt, zip):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 4, 3)).astype(dt)
    p = dataio.write_data_array(tmp_path / "phi", a, zip=zip)
    assert p.name == "phi" + {None: "", "gzip": ".gz", "bzip2": ".bz2"}[zip]
    b = dataio.read_data_array(p, real)
    assert b.dtype == dt and np.array_equal(a, b)


def test_data_array_bytes_are_the_serial_layout(tmp_path):
    a = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4)   # nk=2, nj=3, ni=4
    p = dataio.write_data_array(tmp_path / "d", a)
    raw = p.read_bytes()
    # fwrite(grid_dims_ghostbox, sizeof(int), 3) then fwrite(data, DSZ, num_gridpts), x fastest
    assert raw[:12] == struct.pack("=3i", 4, 3, 2)
    assert raw[12:] == struct.pack("=24d", *range(24))
    g = dataio.write_data_array(tmp_path / "e", a, zip="gzip")
    assert gzip.decompress(g.read_bytes()) == raw


def test_data_array_1d(tmp_path):
    v = np.linspace(0, 1, 7)
    p = dataio.write_data_array_1d(tmp_path / "v", v)
    assert p.read_bytes()[:4] == struct.pack("=i", 7)
    assert np.array_equal(dataio.read_data_array_1d(p), v)


def test_curvature_example_grid_fields_and_round_trip(tmp_path):
    f = dataio.serial_grid(0.04, (-0.75, -0.75, 0.0), (0.75, 0.75, 1.0), "medium")
    assert f["grid_dims"] == [38, 38, 25] and f["grid_dims_ghostbox"] == [44, 44, 31]
    assert f["num_gridpts"] == 44 * 44 * 31
    assert f["gb"] == [0, 43, 0, 43, 0, 30] and f["fb"] == [3, 40, 3, 40, 3, 27]
    assert f["D1_fb"] == [1, 42, 1, 42, 1, 29] and f["D2_fb"] == [2, 41, 2, 41, 2, 28]
    assert f["x_lo_ghostbox"][0] == pytest.approx(-0.87) and f["x_hi"][0] == pytest.approx(0.77)
    assert f["mark_fb"] == 117 and f["beta"] == pytest.approx(0.08) and f["gamma"] == pytest.approx(0.16)
    for real in ("double", "float"):
        p = dataio.write_grid(tmp_path / f"grid_{real}", f, real, zip="gzip")
        g = dataio.read_grid(p, real)
        for k, v in f.items():
            assert g[k] == pytest.approx(v, rel=1e-6), k
    # size: 1 + 3 + 3 ... ints / reals / bytes as written by writeGridToBinaryFile
    p = dataio.write_grid(tmp_path / "g", f, "double")
    n_int = 1 + 3 + 3 + 1 + 5 * 6 + 1
    n_real = 3 * 5 + 2
    assert p.stat().st_size == 4 * n_int + 8 * n_real + 5


@pytest.mark.parametrize("acc,lo,top", [("low", 1, 2), ("medium", 3, 3), ("high", 4, 5), ("very_high", 3, 4)])
def test_fill_boxes_per_accuracy(acc, lo, top):
    f = dataio.serial_grid(0.1, (0.0,) * 3, (1.0,) * 3, acc)
    gd = f["grid_dims_ghostbox"][0]
    assert gd == 10 + 2 * dataio.NUM_GHOSTCELLS[acc]
    assert f["fb"][:2] == [lo, gd - 1 - top]
