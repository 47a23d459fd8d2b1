"""ctypes binding for libraries that export LSMLIB's by-reference entry points.

`F77Library(path)` exposes every symbol of `signatures.SIGNATURES` as a Python
callable.  Arguments follow the reference prototype order with these Python
forms:

  A  a numpy array or a torch tensor (host or device memory) of the
     library's real type; its data pointer is passed
  B  a 6-sequence (ilo, ihi, jlo, jhi, klo, khi)
  R  a Python float
  I  a Python int
  O  omitted from the call; the values are returned (a float, or a tuple
     when the routine has several)
  X  an int32 array (numpy or torch); N the same, in host memory
  U  an int8 or uint8 array on its box (numpy or torch)
  M  a Python int, passed as one byte
  K  omitted from the call; returned like O, as an int

Arrays hold a box `(ilo..ihi, jlo..jhi, klo..khi)` in Fortran order, i.e. a
C-contiguous array of shape (nk, nj, ni) with i fastest.
"""
import ctypes

import numpy as np

from .signatures import SIGNATURES

_REAL_CTYPES = {"double": ctypes.c_double, "float": ctypes.c_float}
_NP_DTYPES = {"double": np.float64, "float": np.float32}


def box_shape(box):
    """(nk, nj, ni) array shape of a box (ilo, ihi, jlo, jhi, klo, khi)."""
    ilo, ihi, jlo, jhi, klo, khi = box
    return (khi - klo + 1, jhi - jlo + 1, ihi - ilo + 1)


def _int_ptr(a, kind):
    """Data pointer of an int32 (kind X/N) or 8-bit (kind U) array."""
    ok = (np.int32,) if kind in "XN" else (np.int8, np.uint8)
    if isinstance(a, np.ndarray):
        if a.dtype.type not in ok:
            raise TypeError(f"kind {kind} array must be {'int32' if kind in 'XN' else 'int8/uint8'}, "
                            f"got {a.dtype}")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        import torch
        want = (torch.int32,) if kind in "XN" else (torch.int8, torch.uint8)
        if a.dtype not in want:
            raise TypeError(f"kind {kind} tensor has dtype {a.dtype}")
        if not a.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return a.data_ptr()
    if a is None:
        return 0
    raise TypeError(f"unsupported array argument {type(a)!r}")


def _data_ptr(a, np_dtype):
    if isinstance(a, np.ndarray):
        if a.dtype != np_dtype:
            raise TypeError(f"array dtype {a.dtype} does not match the library real {np_dtype}")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous (i fastest)")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):  # torch.Tensor
        import torch
        want = torch.float64 if np_dtype == np.float64 else torch.float32
        if a.dtype != want:
            raise TypeError(f"tensor dtype {a.dtype} does not match the library real {want}")
        if not a.is_contiguous():
            raise ValueError("tensors must be contiguous (i fastest)")
        return a.data_ptr()
    if a is None:
        return 0
    raise TypeError(f"unsupported array argument {type(a)!r}")


class F77Library:
    """Bind `path`; symbol names are `prefix + name` for each known routine."""

    def __init__(self, path, real="double", prefix=""):
        if real not in _REAL_CTYPES:
            raise ValueError(real)
        self.path = str(path)
        self.real = real
        self.prefix = prefix
        self.lib = ctypes.CDLL(self.path)
        self._ct = _REAL_CTYPES[real]
        self._np = _NP_DTYPES[real]
        self.dtype = self._np

    def has(self, name):
        try:
            getattr(self.lib, self.prefix + name)
            return True
        except AttributeError:
            return False

    def marshal(self, name, args):
        kinds = SIGNATURES[name]
        cargs, outs, keep = [], [], []
        it = iter(args)
        for k in kinds:
            if k in "OK":
                o = self._ct(0) if k == "O" else ctypes.c_int(0)
                outs.append(o)
                cargs.append(ctypes.byref(o))
                continue
            try:
                a = next(it)
            except StopIteration:
                raise TypeError(f"{name}: too few arguments for kinds {kinds}") from None
            if k == "A":
                cargs.append(ctypes.c_void_p(_data_ptr(a, self._np)))
                keep.append(a)
            elif k == "B":
                b = tuple(int(v) for v in a)
                if len(b) != 6:
                    raise ValueError(f"{name}: a box needs 6 bounds, got {b}")
                for v in b:
                    cargs.append(ctypes.byref(ctypes.c_int(v)))
            elif k == "R":
                cargs.append(ctypes.byref(self._ct(float(a))))
            elif k == "I":
                cargs.append(ctypes.byref(ctypes.c_int(int(a))))
            elif k in "XNU":
                cargs.append(ctypes.c_void_p(_int_ptr(a, k)))
                keep.append(a)
            elif k == "M":
                cargs.append(ctypes.byref(ctypes.c_ubyte(int(a) & 0xFF)))
            else:  # pragma: no cover - table error
                raise ValueError(k)
        rest = list(it)
        if rest:
            raise TypeError(f"{name}: {len(rest)} unexpected extra arguments")
        return cargs, outs, keep

    def call(self, name, *args):
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = None
        cargs, outs, _keep = self.marshal(name, args)
        fn(*cargs)
        if not outs:
            return None
        vals = tuple(o.value for o in outs)
        return vals[0] if len(vals) == 1 else vals

    def __getattr__(self, name):
        if name.startswith("_") or name not in SIGNATURES:
            raise AttributeError(name)
        return lambda *args: self.call(name, *args)
