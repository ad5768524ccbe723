"""HDF5 datasets through libfopis_io.so (include/fopis_io.h): plain
datasets of rank 0 to 3 in C order, paths with intermediate groups
("Potential/phi"), in the element types the reference's files use
(src/Misc/IO.cpp:28-120, RootGridIO.cpp:25-290)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfopis_io.so")

F64, U64, I32, F32, U8 = 0, 1, 2, 3, 4
_TYPE_OF = {np.dtype(np.float64): F64, np.dtype(np.uint64): U64, np.dtype(np.int32): I32,
            np.dtype(np.float32): F32, np.dtype(np.uint8): U8}

_lib = None


def load() -> C.CDLL:
    """The library, loaded once; raises when it was not built (no HDF5 C library)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (the HDF5 C library was not found at build time)")
        L = C.CDLL(LIB_PATH)
        vp, i64p = C.c_void_p, C.POINTER(C.c_int64)
        L.fopis_io_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
        L.fopis_io_close.argtypes = [vp]
        L.fopis_io_write.argtypes = [vp, C.c_char_p, C.c_int, C.c_int, i64p, vp]
        L.fopis_io_exists.argtypes = [vp, C.c_char_p]
        L.fopis_io_info.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), i64p]
        L.fopis_io_read.argtypes = [vp, C.c_char_p, C.c_int, C.c_int64, vp]
        L.fopis_io_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class H5File:
    """An HDF5 file opened for reading (mode "r") or created anew (mode "w")."""

    def __init__(self, path: str, mode: str = "r"):
        if mode not in ("r", "w"):
            raise ValueError("mode must be 'r' or 'w'")
        self.L = load()
        self.h = C.c_void_p()
        self._ok(self.L.fopis_io_open(str(path).encode(), 1 if mode == "w" else 0, C.byref(self.h)))

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.fopis_io_last_error().decode())

    def close(self):
        if self.h:
            self._ok(self.L.fopis_io_close(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write(self, name: str, arr) -> None:
        a = np.array(arr, order="C")  # a 0-d value stays 0-d (a scalar dataspace)
        if a.dtype not in _TYPE_OF:
            raise TypeError(f"no HDF5 element type for {a.dtype}")
        dims = (C.c_int64 * max(1, a.ndim))(*a.shape) if a.ndim else None
        self._ok(self.L.fopis_io_write(self.h, name.encode(), _TYPE_OF[a.dtype], a.ndim, dims,
                                       a.ctypes.data_as(C.c_void_p)))

    def exists(self, name: str) -> bool:
        return self.L.fopis_io_exists(self.h, name.encode()) == 1

    def shape(self, name: str) -> tuple:
        r = C.c_int()
        d = (C.c_int64 * 8)()
        self._ok(self.L.fopis_io_info(self.h, name.encode(), C.byref(r), d))
        return tuple(d[: r.value])

    def read(self, name: str, dtype) -> np.ndarray:
        """The whole dataset, converted to dtype by the HDF5 library."""
        dt = np.dtype(dtype)
        if dt not in _TYPE_OF:
            raise TypeError(f"no HDF5 element type for {dt}")
        out = np.empty(self.shape(name), dt)
        self._ok(self.L.fopis_io_read(self.h, name.encode(), _TYPE_OF[dt], out.size, out.ctypes.data_as(C.c_void_p)))
        return out
