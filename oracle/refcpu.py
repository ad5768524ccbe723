"""ORACLE — test infrastructure only.

ctypes driver for oracle/librefcpu.so, the CPU restatement of the reference
ES step (oracle/refcpu.cpp). Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefcpu.so")


class SpeciesParams(C.Structure):
    _fields_ = [
        ("type", C.c_char * 16),
        ("mass", C.c_double),
        ("charge", C.c_double),
        ("density", C.c_double),
        ("temperature", C.c_double),
        ("drift_kms", C.c_double * 3),
        ("ppc", C.c_int),
        ("injection_axis", C.c_int * 6),
        ("current_density", C.c_double),
    ]


class ObjectParams(C.Structure):
    _fields_ = [
        ("file_name", C.c_char * 512),
        ("is_potential_fixed", C.c_int),
        ("fixed_potential", C.c_double),
        ("initial_potential_offset", C.c_double),
        ("emit_species", C.c_int),
    ]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("dx", C.c_double),
        ("dt", C.c_double),
        ("boundary", C.c_char * 8),
        ("static_b_nT", C.c_double * 3),
        ("shine", C.c_double * 3),
        ("nspecies", C.c_int),
        ("sp", SpeciesParams * 8),
        ("nobjects", C.c_int),
        ("obj", ObjectParams * 4),
        ("post_loop", C.c_int),
    ]


# field selectors (refcpu_capi.cpp fieldPtr)
PHI, RHO, DENSITY, EX, EY, EZ, EXREF, EYREF, EZREF, RESIDUAL, ERROR = range(11)

_libs: dict = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def mutant_path(n: int) -> str:
    """The oracle built with planted mistake n (REFCPU_MUTANT in refcpu.cpp),
    compiled on first use: the pin tests show their checks fail on it."""
    path = os.path.join(HERE, "_mut", f"librefcpu_m{n}.so")
    subprocess.run(["make", "-s", "-C", HERE, path], check=True)
    return path


def lib(path: str | None = None) -> C.CDLL:
    path = path or LIB_PATH
    if path not in _libs:
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(path)
        vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        dp = C.POINTER(C.c_double)
        L.rc_create.restype = vp
        L.rc_create.argtypes = [C.POINTER(Params)]
        L.rc_destroy.argtypes = [vp]
        L.rc_error.restype = C.c_char_p
        for name in ("rc_load_initial_particles", "rc_initialize", "rc_main_loop", "rc_update_velocity",
                     "rc_update_position", "rc_inject", "rc_emit", "rc_update_density", "rc_update_rho",
                     "rc_deposit_only", "rc_solve_poisson", "rc_update_efield", "rc_remove_invalid",
                     "rc_build_cmatrix", "rc_initialize_charge_offset", "rc_set_boundary_phi"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = i
        L.rc_solve_poisson_relres.argtypes = [vp, d, i]
        L.rc_solve_poisson_relres.restype = i
        L.rc_do_psor.argtypes = [vp, i]
        L.rc_do_psor.restype = i
        L.rc_relres.argtypes = [vp]
        L.rc_relres.restype = d
        L.rc_check_phi_residual.argtypes = [vp]
        L.rc_check_phi_residual.restype = d
        L.rc_set_timestep.argtypes = [vp, i]
        L.rc_get_timestep.argtypes = [vp]
        L.rc_get_timestep.restype = i
        for name in ("rc_eps0", "rc_particle_energy", "rc_efield_energy"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = d
        L.rc_species_info.argtypes = [vp, i, dp]
        L.rc_species_info.restype = i
        L.rc_static_b.argtypes = [vp, dp]
        L.rc_calc_flux.argtypes = [vp, i, dp]
        L.rc_calc_flux.restype = i
        L.rc_num_particles.argtypes = [vp, i]
        L.rc_num_particles.restype = i64
        L.rc_get_particles.argtypes = [vp, i] + [dp] * 6 + [C.POINTER(C.c_int32)]
        L.rc_get_particles.restype = i
        L.rc_set_particles.argtypes = [vp, i, i64] + [dp] * 6 + [C.POINTER(C.c_int32)]
        L.rc_set_particles.restype = i
        L.rc_field_size.argtypes = [vp, i, i]
        L.rc_field_size.restype = i64
        L.rc_get_field.argtypes = [vp, i, i, dp]
        L.rc_get_field.restype = i
        L.rc_set_field.argtypes = [vp, i, i, dp]
        L.rc_set_field.restype = i
        L.rc_num_objects.argtypes = [vp]
        L.rc_num_objects.restype = i
        L.rc_object_info.argtypes = [vp, i, dp]
        L.rc_object_info.restype = i
        L.rc_object_cmat_nodes.argtypes = [vp, i, C.POINTER(C.c_int32)]
        L.rc_object_cmat_nodes.restype = i
        L.rc_object_cell_map.argtypes = [vp, i, C.POINTER(C.c_uint8)]
        L.rc_object_cell_map.restype = i
        L.rc_object_cmatrix.argtypes = [vp, i, dp, i]
        L.rc_object_cmatrix.restype = i
        L.rc_object_charge.argtypes = [vp, i, i, dp, i]
        L.rc_object_charge.restype = i
        L.rc_object_redistribute.argtypes = [vp, i]
        L.rc_object_redistribute.restype = i
        L.rc_mt19937_10000th.restype = C.c_uint32
        L.rc_params_size.restype = i64
        L.rc_set_threads.argtypes = [i]
        assert L.rc_params_size() == C.sizeof(Params), "Params layout mismatch"
        _libs[path] = L
    return _libs[path]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def species(type_: str, mass: float, charge: float, density: float, temperature: float, ppc: int,
            drift_kms=(0.0, 0.0, 0.0), injection_axis=(0, 0, 0, 0, 0, 0), current_density: float = 0.0):
    s = SpeciesParams()
    s.type = type_.encode()
    s.mass, s.charge, s.density, s.temperature = mass, charge, density, temperature
    for a in range(3):
        s.drift_kms[a] = drift_kms[a]
    s.ppc = ppc
    for a in range(6):
        s.injection_axis[a] = int(injection_axis[a])
    s.current_density = current_density
    return s


def params(nx, ny, nz, dx, dt, species_list, boundary="DDDDDD", static_b_nT=(0, 0, 0), shine=(0, 0, 0),
           objects=(), post_loop=1000) -> Params:
    p = Params()
    p.nx, p.ny, p.nz, p.dx, p.dt = nx, ny, nz, dx, dt
    p.boundary = boundary.encode()
    for a in range(3):
        p.static_b_nT[a] = static_b_nT[a]
        p.shine[a] = shine[a]
    p.nspecies = len(species_list)
    for q, s in enumerate(species_list):
        p.sp[q] = s
    p.nobjects = len(objects)
    for q, o in enumerate(objects):
        op = ObjectParams()
        op.file_name = o["file_name"].encode()
        op.is_potential_fixed = int(o.get("is_potential_fixed", 0))
        op.fixed_potential = o.get("fixed_potential", 0.0)
        op.initial_potential_offset = o.get("initial_potential_offset", 0.0)
        op.emit_species = o.get("emit_species", -1)
        p.obj[q] = op
    p.post_loop = post_loop
    return p


class Grid:
    """One oracle grid (the reference RootGrid on one rank)."""

    def __init__(self, p: Params, lib_path: str | None = None):
        self.L = lib(lib_path)
        self.p = p
        self.h = self.L.rc_create(C.byref(p))
        if not self.h:
            raise RuntimeError(self.L.rc_error().decode())
        self.nx, self.ny, self.nz = p.nx, p.ny, p.nz
        self.S = p.nspecies

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rc_destroy(self.h)
            self.h = None

    def _ok(self, rc):
        if rc < 0:
            raise RuntimeError(self.L.rc_error().decode())
        return rc

    def call(self, name, *args):
        return self._ok(getattr(self.L, name)(self.h, *args))

    # ---- particles
    def num_particles(self, s: int) -> int:
        return int(self.L.rc_num_particles(self.h, s))

    def get_particles(self, s: int) -> dict:
        n = self.num_particles(s)
        arrs = [np.empty(n, np.float64) for _ in range(6)]
        valid = np.empty(n, np.int32)
        self._ok(self.L.rc_get_particles(self.h, s, *[_dp(a) for a in arrs],
                                         valid.ctypes.data_as(C.POINTER(C.c_int32))))
        out = dict(zip(("x", "y", "z", "vx", "vy", "vz"), arrs))
        out["valid"] = valid
        return out

    def set_particles(self, s: int, d: dict, valid=None) -> None:
        arrs = [np.ascontiguousarray(d[k], dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")]
        n = len(arrs[0])
        vp = None
        if valid is not None:
            valid = np.ascontiguousarray(valid, dtype=np.int32)
            vp = valid.ctypes.data_as(C.POINTER(C.c_int32))
        self._ok(self.L.rc_set_particles(self.h, s, n, *[_dp(a) for a in arrs], vp))

    # ---- fields
    def field_shape(self, which: int):
        cx, cy, cz = self.nx + 2, self.ny + 2, self.nz + 2
        return {PHI: (cx, cy, cz), RHO: (cx, cy, cz), DENSITY: (cx - 1, cy - 1, cz - 1),
                EX: (cx - 1, cy, cz), EY: (cx, cy - 1, cz), EZ: (cx, cy, cz - 1), EXREF: (cx, cy, cz),
                EYREF: (cx, cy, cz), EZREF: (cx, cy, cz), RESIDUAL: (cx, cy, cz), ERROR: (cx, cy, cz)}[which]

    def get_field(self, which: int, s: int = 0) -> np.ndarray:
        out = np.empty(self.field_shape(which), np.float64)
        self._ok(self.L.rc_get_field(self.h, which, s, _dp(out)))
        return out

    def set_field(self, which: int, arr: np.ndarray, s: int = 0) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.field_shape(which))
        self._ok(self.L.rc_set_field(self.h, which, s, _dp(a)))

    def eps0(self) -> float:
        return float(self.L.rc_eps0(self.h))

    def species_info(self, s: int) -> dict:
        out = np.zeros(9)
        self._ok(self.L.rc_species_info(self.h, s, _dp(out)))
        return {"charge": out[0], "mass": out[1], "size": out[2], "q": out[3], "vth": out[4],
                "deviation": out[5], "drift": out[6:9].copy()}

    def static_b(self) -> np.ndarray:
        out = np.zeros(3)
        self.L.rc_static_b(self.h, _dp(out))
        return out

    def calc_flux(self, s: int) -> np.ndarray:
        out = np.zeros(6)
        self._ok(self.L.rc_calc_flux(self.h, s, _dp(out)))
        return out

    # ---- objects
    def object_info(self, o: int = 0) -> dict:
        out = np.zeros(5)
        self._ok(self.L.rc_object_info(self.h, o, _dp(out)))
        return {"ncmat": int(out[0]), "potential": out[1], "total_charge": out[2], "total_cmat": out[3],
                "defined": bool(out[4])}

    def cmat_nodes(self, o: int = 0) -> np.ndarray:
        n = self.object_info(o)["ncmat"]
        out = np.zeros((n, 3), np.int32)
        self._ok(self.L.rc_object_cmat_nodes(self.h, o, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def cell_map(self, o: int = 0) -> np.ndarray:
        out = np.zeros((self.nx + 1, self.ny + 1, self.nz + 1), np.uint8)
        self._ok(self.L.rc_object_cell_map(self.h, o, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def cmatrix(self, o: int = 0) -> np.ndarray:
        n = self.object_info(o)["ncmat"]
        out = np.zeros((n, n))
        self._ok(self.L.rc_object_cmatrix(self.h, o, _dp(out), 0))
        return out

    def set_cmatrix(self, C_: np.ndarray, o: int = 0) -> None:
        a = np.ascontiguousarray(C_, dtype=np.float64)
        self._ok(self.L.rc_object_cmatrix(self.h, o, _dp(a), 1))

    def object_charge(self, which: int, o: int = 0) -> np.ndarray:
        n = self.object_info(o)["ncmat"]
        size = self.S if which == 2 else self.S * n
        out = np.zeros(size)
        self._ok(self.L.rc_object_charge(self.h, o, which, _dp(out), 0))
        return out if which == 2 else out.reshape(self.S, n)

    def set_object_charge(self, which: int, arr: np.ndarray, o: int = 0) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._ok(self.L.rc_object_charge(self.h, o, which, _dp(a), 1))


def mt19937_10000th() -> int:
    return int(lib().rc_mt19937_10000th())


def set_threads(n: int) -> None:
    lib().rc_set_threads(int(n))
