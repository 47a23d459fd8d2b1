"""LSMLIB serial-package data files, for exchanging fields with its users.

Formats (native byte order, as the serial package writes them with fwrite):

  data array   (writeDataArray / readDataArray, src/serial/lsm_data_arrays.c:363-436)
               int32 grid_dims_ghostbox[3] (x, y, z), then the ghost-box values,
               x fastest, as LSMLIB_REAL (8 bytes in the double build, 4 in float)
  1-d array    (writeDataArray1d / readDataArray1d, :439-490)
               int32 num_elements, then the values
  grid         (writeGridToBinaryFile / readGridFromBinaryFile, lsm_grid.c:604-760)
               the Grid structure field by field (see GRID_FIELDS)

A name ending in .gz (.bz2) is read through gzip (bzip2), as checkUnzipFile
(lsm_file.c:18-60) does by running gunzip; zip="gzip" or "bzip2" writes the
compressed file `path + ".gz"` (".bz2"), as zipFile runs gzip -f on it.
Arrays are numpy (nk, nj, ni) (the library's layout, i fastest); torch
tensors are accepted for writing.
"""
import bz2
import gzip
from pathlib import Path

import numpy as np

_REAL = {"double": np.float64, "float": np.float32}

# (name, kind, count) in the order of writeGridToBinaryFile (lsm_grid.c:617-672);
# kind: i = int32, r = LSMLIB_REAL, b = unsigned char
GRID_FIELDS = [
    ("num_dims", "i", 1), ("x_lo", "r", 3), ("x_hi", "r", 3), ("x_lo_ghostbox", "r", 3),
    ("x_hi_ghostbox", "r", 3), ("grid_dims", "i", 3), ("grid_dims_ghostbox", "i", 3), ("dx", "r", 3),
    ("num_gridpts", "i", 1), ("gb", "i", 6), ("fb", "i", 6), ("D1_fb", "i", 6), ("D2_fb", "i", 6),
    ("D3_fb", "i", 6), ("num_nb_levels", "i", 1), ("mark_gb", "b", 1), ("mark_D1", "b", 1),
    ("mark_D2", "b", 1), ("mark_D3", "b", 1), ("mark_fb", "b", 1), ("beta", "r", 1), ("gamma", "r", 1),
]

# ghost cells per accuracy (LOW, MEDIUM, HIGH, VERY_HIGH), lsm_grid.c:46
NUM_GHOSTCELLS = {"low": 2, "medium": 3, "high": 5, "very_high": 4}


def _open_read(path):
    p = str(path)
    if ".gz" in p:
        return gzip.open(p, "rb")
    if ".bz2" in p:
        return bz2.open(p, "rb")
    return open(p, "rb")


def _write(path, payload, zip=None):
    p = Path(path)
    if zip is None:
        p.write_bytes(payload)
        return p
    if zip == "gzip":
        q = Path(str(p) + ".gz")
        with gzip.open(q, "wb") as f:
            f.write(payload)
        return q
    if zip == "bzip2":
        q = Path(str(p) + ".bz2")
        with bz2.open(q, "wb") as f:
            f.write(payload)
        return q
    raise ValueError(f"unknown zip {zip!r}")


def _as_numpy(data):
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    return np.ascontiguousarray(data)


def write_data_array(path, data, zip=None):
    """writeDataArray: `data` (nk, nj, ni) over the ghost box, float64 or
    float32 (the LSMLIB_REAL of the file).  Returns the path written."""
    a = _as_numpy(data)
    if a.ndim != 3 or a.dtype not in (np.float64, np.float32):
        raise ValueError("data must be a 3-d float64 or float32 array (nk, nj, ni)")
    nk, nj, ni = a.shape
    head = np.array([ni, nj, nk], dtype=np.int32).tobytes()
    return _write(path, head + a.tobytes(), zip)


def read_data_array(path, real="double"):
    """readDataArray: returns the (nk, nj, ni) array of the file."""
    with _open_read(path) as f:
        raw = f.read()
    dims = np.frombuffer(raw[:12], dtype=np.int32)
    ni, nj, nk = (int(v) for v in dims)
    dt = _REAL[real]
    n = ni * nj * nk
    vals = np.frombuffer(raw[12:12 + n * np.dtype(dt).itemsize], dtype=dt)
    if vals.size != n:
        raise ValueError(f"{path}: {vals.size} values for a {ni} x {nj} x {nk} ghost box")
    return vals.reshape(nk, nj, ni).copy()


def write_data_array_1d(path, data, zip=None):
    """writeDataArray1d."""
    a = _as_numpy(data).ravel()
    return _write(path, np.array([a.size], dtype=np.int32).tobytes() + a.tobytes(), zip)


def read_data_array_1d(path, real="double"):
    """readDataArray1d."""
    with _open_read(path) as f:
        raw = f.read()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    return np.frombuffer(raw[4:4 + n * np.dtype(_REAL[real]).itemsize], dtype=_REAL[real]).copy()


def serial_grid(dx, x_lo, x_hi, accuracy="medium", num_dims=3):
    """The Grid fields createGridSetDx builds (lsm_grid.c:50-114) with the
    index-space limits of setIndexSpaceLimits (:986-1030, :770-980): every
    index space is 0-based over the ghost box."""
    g = NUM_GHOSTCELLS[accuracy]
    dims = [int(np.rint((x_hi[d] - x_lo[d]) / dx)) for d in range(3)]
    gd = [n + 2 * g for n in dims]
    f = {"num_dims": num_dims, "x_lo": list(x_lo), "x_hi": [x_lo[d] + dims[d] * dx for d in range(3)],
         "grid_dims": dims, "grid_dims_ghostbox": gd, "dx": [dx] * 3}
    f["x_lo_ghostbox"] = [f["x_lo"][d] - g * dx for d in range(3)]
    f["x_hi_ghostbox"] = [f["x_hi"][d] + g * dx for d in range(3)]
    f["num_gridpts"] = gd[0] * gd[1] * gd[2]
    lim = lambda lo: [v for d in range(3) for v in (lo, gd[d] - 1 - lo)]  # noqa: E731
    f["gb"], f["D1_fb"], f["D2_fb"], f["D3_fb"] = lim(0), lim(1), lim(2), lim(3)
    # the common fill box of the plus and minus derivatives (lsm_grid.c:815-980);
    # (lower index, cells below the top of the ghost box)
    lo, top = {"low": (1, 2), "medium": (3, 3), "high": (4, 5), "very_high": (3, 4)}[accuracy]
    f["fb"] = [v for d in range(3) for v in (lo, gd[d] - 1 - top)]
    f["num_nb_levels"] = g
    f["mark_gb"] = 120
    f["mark_D1"], f["mark_D2"], f["mark_D3"] = 119, 118, 117
    f["mark_fb"] = 120 - g
    big = accuracy in ("high", "very_high")
    f["beta"] = (3 if big else 2) * dx
    f["gamma"] = (6 if big else 4) * dx
    return f


def write_grid(path, fields, real="double", zip=None):
    """writeGridToBinaryFile of a dict with the keys of GRID_FIELDS."""
    dt = _REAL[real]
    out = b""
    for name, kind, count in GRID_FIELDS:
        v = np.atleast_1d(np.asarray(fields[name]))
        if v.size != count:
            raise ValueError(f"grid field {name}: {v.size} values, want {count}")
        out += v.astype({"i": np.int32, "r": dt, "b": np.uint8}[kind]).tobytes()
    return _write(path, out, zip)


def read_grid(path, real="double"):
    """readGridFromBinaryFile: a dict with the keys of GRID_FIELDS."""
    dt = _REAL[real]
    with _open_read(path) as f:
        raw = f.read()
    pos, out = 0, {}
    for name, kind, count in GRID_FIELDS:
        t = np.dtype({"i": np.int32, "r": dt, "b": np.uint8}[kind])
        v = np.frombuffer(raw[pos:pos + count * t.itemsize], dtype=t)
        pos += count * t.itemsize
        out[name] = v.tolist() if count > 1 else v[0].item()
    return out
