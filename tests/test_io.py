"""libfopis_io.so (include/fopis_io.h): HDF5 datasets in the layouts the
reference keeps between runs, checked on the CPU.

Round trips of every element type and rank; missing datasets and shape
mismatches are errors; and h5dump (the HDF5 tools' own reader) sees the
reference's types and dataspaces: a single double as a SCALAR
H5T_IEEE_F64LE (HighFive::DataSpace::From(double)), size_t vectors as
H5T_STD_U64LE, int as H5T_STD_I32LE, matrices as SIMPLE {(n, n)}
(src/Misc/IO.cpp:28-57, RootGridIO.cpp:133-181, Environment.cpp:421-438)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fopis_amd", "libfopis_io.so")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libfopis_io.so not built (no HDF5)")

F64, U64, I32, F32, U8 = 0, 1, 2, 3, 4


class IO:
    def __init__(self):
        L = C.CDLL(LIB)
        vp, i64p = C.c_void_p, C.POINTER(C.c_int64)
        L.fopis_io_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
        L.fopis_io_close.argtypes = [vp]
        L.fopis_io_write.argtypes = [vp, C.c_char_p, C.c_int, C.c_int, i64p, vp]
        L.fopis_io_exists.argtypes = [vp, C.c_char_p]
        L.fopis_io_info.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), i64p]
        L.fopis_io_read.argtypes = [vp, C.c_char_p, C.c_int, C.c_int64, vp]
        L.fopis_io_last_error.restype = C.c_char_p
        self.L = L

    def open(self, path, mode):
        h = C.c_void_p()
        rc = self.L.fopis_io_open(str(path).encode(), mode, C.byref(h))
        if rc:
            raise RuntimeError(self.L.fopis_io_last_error().decode())
        return h

    def write(self, h, name, arr, typ):
        a = np.array(arr, order="C")  # keeps a 0-d value 0-d (a scalar dataspace)
        dims = (C.c_int64 * max(1, a.ndim))(*a.shape) if a.ndim else None
        rc = self.L.fopis_io_write(h, name.encode(), typ, a.ndim, dims, a.ctypes.data_as(C.c_void_p))
        if rc:
            raise RuntimeError(self.L.fopis_io_last_error().decode())

    def info(self, h, name):
        r = C.c_int()
        d = (C.c_int64 * 8)()
        if self.L.fopis_io_info(h, name.encode(), C.byref(r), d):
            raise KeyError(self.L.fopis_io_last_error().decode())
        return tuple(d[: r.value])

    def read(self, h, name, dtype, typ):
        shape = self.info(h, name)
        out = np.empty(shape, dtype)
        if self.L.fopis_io_read(h, name.encode(), typ, out.size, out.ctypes.data_as(C.c_void_p)):
            raise RuntimeError(self.L.fopis_io_last_error().decode())
        return out


def test_round_trip_types_and_ranks(tmp_path):
    io = IO()
    rng = np.random.default_rng(3)
    p = tmp_path / "a.h5"
    h = io.open(p, 1)
    data = {
        "grid_spacing": (np.float64(2.35), F64, np.float64),
        "timestep": (np.int32(501), I32, np.int32),
        "electron/generated_counts": (np.array([7, 7, 7, 5, 5, 5], np.uint64), U64, np.uint64),
        "electron/x": (rng.random(1000), F64, np.float64),
        "capacity_matrix": (rng.random((17, 17)), F64, np.float64),
        "Potential/phi": (rng.random((5, 6, 7)), F64, np.float64),
        "Bfield/bx": (np.zeros((0, 0, 0)), F64, np.float64),
        "pairs/000000/rho_hat": (rng.standard_normal((4, 5, 6)).astype(np.float32), F32, np.float32),
        "pairs/000000/mask": (rng.integers(0, 3, (4, 5, 6)).astype(np.uint8), U8, np.uint8),
    }
    for name, (a, t, _) in data.items():
        io.write(h, name, a, t)
    with pytest.raises(RuntimeError):
        io.write(h, "electron/x", data["electron/x"][0], F64)  # no overwrite
    io.L.fopis_io_close(h)
    h = io.open(p, 0)
    for name, (a, t, dt) in data.items():
        got = io.read(h, name, dt, t)
        assert got.shape == np.shape(a)
        np.testing.assert_array_equal(got, a)
    assert io.L.fopis_io_exists(h, b"Potential/phi") == 1
    assert io.L.fopis_io_exists(h, b"Potential/rho") == 0
    with pytest.raises(KeyError):
        io.info(h, "ion/x")
    bad = np.empty(3)
    assert io.L.fopis_io_read(h, b"electron/x", F64, 3, bad.ctypes.data_as(C.c_void_p)) != 0  # wrong count
    # conversion on read: the stored int as a double
    t = io.read(h, "timestep", np.float64, F64)
    assert t == 501.0
    io.L.fopis_io_close(h)
    with pytest.raises(RuntimeError):
        io.open(tmp_path / "missing.h5", 0)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_h5dump_sees_the_reference_layout(tmp_path):
    io = IO()
    p = tmp_path / "cube16.obj.h5"
    h = io.open(p, 1)
    io.write(h, "capacity_matrix", np.eye(4), F64)
    io.write(h, "grid_spacing", np.float64(0.3), F64)
    io.write(h, "time_spacing", np.float64(1e-7), F64)
    io.write(h, "electron/generated_counts", np.zeros(6, np.uint64), U64)
    io.write(h, "timestep", np.int32(3), I32)
    io.write(h, "pairs/000000/rho_hat", np.zeros((2, 2, 2), np.float32), F32)
    io.write(h, "pairs/000000/mask", np.zeros((2, 2, 2), np.uint8), U8)
    io.L.fopis_io_close(h)
    out = subprocess.run([H5DUMP, "-H", str(p)], capture_output=True, text=True, check=True).stdout
    blocks = dict(re.findall(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+)', out))
    assert blocks == {"capacity_matrix": "H5T_IEEE_F64LE", "grid_spacing": "H5T_IEEE_F64LE",
                      "time_spacing": "H5T_IEEE_F64LE", "generated_counts": "H5T_STD_U64LE",
                      "timestep": "H5T_STD_I32LE", "rho_hat": "H5T_IEEE_F32LE", "mask": "H5T_STD_U8LE"}
    assert re.search(r'DATASET "capacity_matrix".*?DATASPACE\s+SIMPLE \{ \( 4, 4 \)', out, re.S)
    assert re.search(r'DATASET "grid_spacing".*?DATASPACE\s+SCALAR', out, re.S)
    assert 'GROUP "electron"' in out


def test_package_binding_round_trip(tmp_path):
    """fopis_amd.h5 (the package's binding) picks the element type from the
    array's dtype and reads back bit for bit; a missing dataset raises."""
    from fopis_amd.h5 import H5File

    rng = np.random.default_rng(8)
    data = {"grid": np.array([8, 16, 24], np.int32), "count": np.int32(2),
            "pairs/000001/phi_hat": rng.standard_normal((8, 16, 24)).astype(np.float32),
            "pairs/000001/mask": (rng.random((8, 16, 24)) < 0.1).astype(np.uint8),
            "capacity_matrix": rng.random((5, 5)), "generated": np.arange(6, dtype=np.uint64)}
    p = tmp_path / "b.h5"
    with H5File(p, "w") as f:
        for k, v in data.items():
            f.write(k, v)
    with H5File(p, "r") as f:
        for k, v in data.items():
            got = f.read(k, np.asarray(v).dtype)
            assert got.shape == np.shape(v)
            np.testing.assert_array_equal(got, v)
        assert f.exists("pairs/000001/mask") and not f.exists("pairs/000002/mask")
        with pytest.raises(RuntimeError):
            f.read("pairs/000002/mask", np.uint8)
