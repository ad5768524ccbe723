"""ctypes binding of libfopis_hip.so (include/fopis.h).

This is the Python mirror of the C ABI used by the tests and the benchmark.
It loads the in-tree library and raises if it is missing: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfopis_hip.so")

FP64, FP32 = 0, 1
PHI, RHO, DENSITY, EX, EY, EZ, EXREF, EYREF, EZREF = range(9)
MAX_SPECIES = 8

EXPORTS = [
    "fopis_create", "fopis_destroy", "fopis_last_error", "fopis_device_info", "fopis_set_particles",
    "fopis_num_particles", "fopis_get_particles", "fopis_add_particles", "fopis_sort_by_cell",
    "fopis_get_cell_keys", "fopis_update_particle_velocity_es", "fopis_update_particle_position_es",
    "fopis_update_density", "fopis_update_rho", "fopis_solve_poisson", "fopis_update_efield",
    "fopis_main_loop_es", "fopis_initialize", "fopis_get_field", "fopis_set_field", "fopis_energy",
    "fopis_set_object", "fopis_build_cmatrix", "fopis_get_cmatrix", "fopis_set_cmatrix",
    "fopis_add_object_charge", "fopis_get_object_state", "fopis_set_object_charge_map", "fopis_phase_time",
    "fopis_reset_timers", "fopis_synchronize", "fopis_solver_stats", "fopis_load_maxwellian",
    "fopis_guess_init", "fopis_guess_num_params", "fopis_guess_get_params", "fopis_guess_set_params",
    "fopis_guess_set_mode", "fopis_guess_apply", "fopis_guess_dataset_alloc", "fopis_guess_dataset_store",
    "fopis_guess_dataset_get", "fopis_guess_dataset_set", "fopis_guess_grad", "fopis_guess_train",
    "fopis_guess_eval", "fopis_guess_synth_boxes", "fopis_nn_conv3d", "fopis_nn_conv3d_backward",
    "fopis_first_solve_cycles", "fopis_moments", "fopis_guess_set_loss", "fopis_inject_maxwellian",
    "fopis_emit_photoelectrons", "fopis_particle_ledger", "fopis_solver_passes",
]


class SpeciesDesc(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("size", C.c_double)]


class Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("boundary", C.c_char * 8),
        ("nspecies", C.c_int32),
        ("species", SpeciesDesc * MAX_SPECIES),
        ("static_b", C.c_double * 3),
        ("eps0", C.c_double),
        ("precision", C.c_int32),
        ("mg_precision", C.c_int32),
        ("mg_omega", C.c_double),
        ("mg_pre", C.c_int32), ("mg_post", C.c_int32),
        ("mg_max_cycles", C.c_int32),
        ("poisson_tol", C.c_double),
        ("capacity_slack", C.c_double),
    ]


class FopisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FopisError(-3, f"{LIB_PATH} is missing: build it with python fopis_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    dp = C.POINTER(C.c_double)
    L.fopis_last_error.restype = C.c_char_p
    L.fopis_device_info.argtypes = [C.c_char_p, C.c_int]
    L.fopis_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    L.fopis_destroy.argtypes = [vp]
    L.fopis_set_particles.argtypes = [vp, C.c_int, i64] + [dp] * 6
    L.fopis_add_particles.argtypes = [vp, C.c_int, i64] + [dp] * 6
    L.fopis_num_particles.argtypes = [vp, C.c_int, C.POINTER(i64)]
    L.fopis_get_particles.argtypes = [vp, C.c_int, i64] + [dp] * 6 + [C.POINTER(i64)]
    L.fopis_get_cell_keys.argtypes = [vp, C.c_int, i64, C.POINTER(C.c_int32), C.POINTER(i64)]
    for name in ("fopis_sort_by_cell", "fopis_update_particle_velocity_es", "fopis_update_particle_position_es",
                 "fopis_update_density", "fopis_update_rho", "fopis_update_efield", "fopis_main_loop_es",
                 "fopis_initialize", "fopis_reset_timers", "fopis_synchronize"):
        getattr(L, name).argtypes = [vp]
    L.fopis_solve_poisson.argtypes = [vp, C.POINTER(i32), dp]
    L.fopis_get_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.fopis_set_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.fopis_energy.argtypes = [vp, dp]
    L.fopis_set_object.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8), i32, C.POINTER(C.c_int32), i32, d]
    L.fopis_build_cmatrix.argtypes = [vp, C.c_int]
    L.fopis_get_cmatrix.argtypes = [vp, C.c_int, dp]
    L.fopis_set_cmatrix.argtypes = [vp, C.c_int, dp]
    L.fopis_add_object_charge.argtypes = [vp, C.c_int, dp, dp]
    L.fopis_get_object_state.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.fopis_set_object_charge_map.argtypes = [vp, C.c_int, dp]
    L.fopis_phase_time.argtypes = [vp, C.c_char_p, dp, C.POINTER(i64)]
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_info() -> str:
    L = load()
    buf = C.create_string_buffer(256)
    rc = L.fopis_device_info(buf, 256)
    if rc != 0:
        raise FopisError(rc, L.fopis_last_error().decode())
    return buf.value.decode()


def make_desc(nx, ny, nz, species, eps0, static_b=(0.0, 0.0, 0.0), boundary="DDDDDD", precision=FP64,
              mg_precision=None, mg_omega=1.25, mg_pre=2, mg_post=2, mg_max_cycles=100, poisson_tol=1e-8,
              capacity_slack=1.25) -> Desc:
    """species: list of (charge, mass, size) in the reference's normalised units."""
    d = Desc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.boundary = boundary.encode()
    d.nspecies = len(species)
    for q, (ch, m, sz) in enumerate(species):
        d.species[q] = SpeciesDesc(ch, m, sz)
    for a in range(3):
        d.static_b[a] = static_b[a]
    d.eps0 = eps0
    d.precision = precision
    d.mg_precision = precision if mg_precision is None else mg_precision
    d.mg_omega, d.mg_pre, d.mg_post = mg_omega, mg_pre, mg_post
    d.mg_max_cycles = mg_max_cycles
    d.poisson_tol = poisson_tol
    d.capacity_slack = capacity_slack
    return d


class Context:
    """One device context: fields, tile-binned particles, conductors."""

    def __init__(self, desc: Desc):
        self.L = load()
        self.desc = desc
        h = C.c_void_p()
        self._ok(self.L.fopis_create(C.byref(desc), C.byref(h)))
        self.h = h
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz
        self.S = desc.nspecies

    def close(self):
        if getattr(self, "h", None):
            self.L.fopis_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, rc):
        if rc != 0:
            raise FopisError(rc, self.L.fopis_last_error().decode())
        return rc

    # ---- particles
    def set_particles(self, s: int, d: dict) -> None:
        arrs = [np.ascontiguousarray(d[k], dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")]
        self._ok(self.L.fopis_set_particles(self.h, s, len(arrs[0]), *[_dp(a) for a in arrs]))

    def add_particles(self, s: int, d: dict) -> None:
        arrs = [np.ascontiguousarray(d[k], dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")]
        self._ok(self.L.fopis_add_particles(self.h, s, len(arrs[0]), *[_dp(a) for a in arrs]))

    def num_particles(self, s: int) -> int:
        n = C.c_int64()
        self._ok(self.L.fopis_num_particles(self.h, s, C.byref(n)))
        return n.value

    def get_particles(self, s: int) -> dict:
        cap = self.num_particles(s)
        arrs = [np.empty(cap, np.float64) for _ in range(6)]
        n = C.c_int64()
        self._ok(self.L.fopis_get_particles(self.h, s, cap, *[_dp(a) for a in arrs], C.byref(n)))
        return dict(zip(("x", "y", "z", "vx", "vy", "vz"), [a[: n.value] for a in arrs]))

    def load_maxwellian(self, s: int, n: int, seed: int, lo, hi, drift, sigma: float) -> None:
        """Replace species s by n particles drawn on the device (uniform in [lo, hi), N(drift, sigma))."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (lo, hi, drift)]
        self.L.fopis_load_maxwellian.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double]
        self._ok(self.L.fopis_load_maxwellian(self.h, s, n, seed, *[_dp(v) for v in a], sigma))

    def cell_keys(self, s: int) -> np.ndarray:
        cap = self.num_particles(s)
        out = np.empty(cap, np.int32)
        n = C.c_int64()
        self._ok(self.L.fopis_get_cell_keys(self.h, s, cap, out.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        return out[: n.value]

    def sort_by_cell(self) -> None:
        self._ok(self.L.fopis_sort_by_cell(self.h))

    # ---- step stages
    def update_particle_velocity_es(self):
        self._ok(self.L.fopis_update_particle_velocity_es(self.h))

    def update_particle_position_es(self):
        self._ok(self.L.fopis_update_particle_position_es(self.h))

    def update_density(self):
        self._ok(self.L.fopis_update_density(self.h))

    def update_rho(self):
        self._ok(self.L.fopis_update_rho(self.h))

    def solve_poisson(self):
        cyc = C.c_int32()
        rr = C.c_double()
        self._ok(self.L.fopis_solve_poisson(self.h, C.byref(cyc), C.byref(rr)))
        return cyc.value, rr.value

    def update_efield(self):
        self._ok(self.L.fopis_update_efield(self.h))

    def initialize(self):
        self._ok(self.L.fopis_initialize(self.h))

    def main_loop_es(self):
        self._ok(self.L.fopis_main_loop_es(self.h))

    def synchronize(self):
        self._ok(self.L.fopis_synchronize(self.h))

    # ---- fields
    def field_shape(self, which: int):
        cx, cy, cz = self.nx + 2, self.ny + 2, self.nz + 2
        return {PHI: (cx, cy, cz), RHO: (cx, cy, cz), DENSITY: (cx - 1, cy - 1, cz - 1), EX: (cx - 1, cy, cz),
                EY: (cx, cy - 1, cz), EZ: (cx, cy, cz - 1), EXREF: (cx, cy, cz), EYREF: (cx, cy, cz),
                EZREF: (cx, cy, cz)}[which]

    def get_field(self, which: int, s: int = 0) -> np.ndarray:
        out = np.empty(self.field_shape(which), np.float64)
        self._ok(self.L.fopis_get_field(self.h, which, s, _dp(out)))
        return out

    def set_field(self, which: int, arr: np.ndarray, s: int = 0) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.field_shape(which))
        self._ok(self.L.fopis_set_field(self.h, which, s, _dp(a)))

    def inject_maxwellian(self, s: int, counts, seed: int, drift, vth: float, sigma: float) -> None:
        """Boundary injection drawn on the device: counts[face] particles per face (x-, x+, y-, y+, z-, z+)."""
        c6 = np.ascontiguousarray(counts, dtype=np.int64)
        dr = np.ascontiguousarray(drift, dtype=np.float64)
        self.L.fopis_inject_maxwellian.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.c_uint64,
                                                   C.POINTER(C.c_double), C.c_double, C.c_double]
        self._ok(self.L.fopis_inject_maxwellian(self.h, s, c6.ctypes.data_as(C.POINTER(C.c_int64)), seed, _dp(dr),
                                                vth, sigma))

    def emit_photoelectrons(self, o: int, s: int, shine, amount: float, vth: float, seed: int) -> None:
        """Photoemission from object o's lit face drawn on the device (species s, ceil(amount) per node)."""
        sh = np.ascontiguousarray(shine, dtype=np.float64)
        self.L.fopis_emit_photoelectrons.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_double,
                                                     C.c_double, C.c_uint64]
        self._ok(self.L.fopis_emit_photoelectrons(self.h, o, s, _dp(sh), amount, vth, seed))

    def moments(self, s: int) -> dict:
        """count, mean position, mean velocity and velocity variance of species s (FP64)."""
        out = np.zeros(10)
        self.L.fopis_moments.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        self._ok(self.L.fopis_moments(self.h, s, _dp(out)))
        return {"count": out[0], "mean_x": out[1:4], "mean_v": out[4:7], "var_v": out[7:10]}

    def ledger(self, s: int) -> dict:
        """Species s since the context was made: staged particles that joined,
        particles absorbed at a domain face, particles captured by a conductor."""
        out = np.zeros(3, np.int64)
        self.L.fopis_particle_ledger.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        self._ok(self.L.fopis_particle_ledger(self.h, s, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return {"joined": int(out[0]), "absorbed": int(out[1]), "captured": int(out[2])}

    def energy(self):
        out = np.zeros(2)
        self._ok(self.L.fopis_energy(self.h, _dp(out)))
        return out[0], out[1]

    # ---- conductors
    def set_object(self, o: int, cell_mask: np.ndarray, cmat_ijk: np.ndarray, fixed=False, fixed_potential=0.0):
        m = np.ascontiguousarray(cell_mask, dtype=np.uint8)
        ijk = np.ascontiguousarray(cmat_ijk, dtype=np.int32).reshape(-1, 3)
        self._ok(self.L.fopis_set_object(self.h, o, m.ctypes.data_as(C.POINTER(C.c_uint8)), len(ijk),
                                         ijk.ctypes.data_as(C.POINTER(C.c_int32)), int(fixed), fixed_potential))
        self._ncmat = getattr(self, "_ncmat", {})
        self._ncmat[o] = len(ijk)

    def build_cmatrix(self, o: int = 0):
        self._ok(self.L.fopis_build_cmatrix(self.h, o))

    def get_cmatrix(self, o: int = 0) -> np.ndarray:
        n = self._ncmat[o]
        out = np.zeros((n, n))
        self._ok(self.L.fopis_get_cmatrix(self.h, o, _dp(out)))
        return out

    def set_cmatrix(self, Cm: np.ndarray, o: int = 0):
        a = np.ascontiguousarray(Cm, dtype=np.float64)
        self._ok(self.L.fopis_set_cmatrix(self.h, o, _dp(a)))

    def add_object_charge(self, o: int, dq=None, dcur=None):
        a = None if dq is None else np.ascontiguousarray(dq, dtype=np.float64).ravel()
        b = None if dcur is None else np.ascontiguousarray(dcur, dtype=np.float64).ravel()
        self._ok(self.L.fopis_add_object_charge(self.h, o, _dp(a) if a is not None else None,
                                                _dp(b) if b is not None else None))

    def object_state(self, o: int = 0) -> dict:
        n = self._ncmat[o]
        pot, q = C.c_double(), C.c_double()
        cur = np.zeros(self.S)
        cmap = np.zeros(self.S * max(n, 1))
        self._ok(self.L.fopis_get_object_state(self.h, o, C.byref(pot), C.byref(q), _dp(cur), _dp(cmap)))
        lost = np.zeros(self.S, np.int64)
        self.L.fopis_object_unattributed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        self._ok(self.L.fopis_object_unattributed(self.h, o, lost.ctypes.data_as(C.POINTER(C.c_int64))))
        return {"potential": pot.value, "total_charge": q.value, "current": cur,
                "charge_map": cmap[: self.S * n].reshape(self.S, n), "unattributed": lost}

    def set_object_charge_map(self, o: int, cmap: np.ndarray):
        a = np.ascontiguousarray(cmap, dtype=np.float64).ravel()
        self._ok(self.L.fopis_set_object_charge_map(self.h, o, _dp(a)))

    # ---- timing
    def phase_time(self, phase: str):
        avg, n = C.c_double(), C.c_int64()
        self._ok(self.L.fopis_phase_time(self.h, phase.encode(), C.byref(avg), C.byref(n)))
        return avg.value, n.value

    def reset_timers(self):
        self._ok(self.L.fopis_reset_timers(self.h))


def _bind_stats(L):
    if getattr(L, "_stats_bound", False):
        return
    L.fopis_solver_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L._stats_bound = True


def solver_stats(ctx: "Context") -> dict:
    """Poisson cycles/relres of the last solve and totals since reset_timers
    (first_cycles: summed over the first solve of each step)."""
    _bind_stats(ctx.L)
    lc, lr, tc, ts, fc = C.c_int32(), C.c_double(), C.c_int64(), C.c_int64(), C.c_int64()
    ctx._ok(ctx.L.fopis_solver_stats(ctx.h, C.byref(lc), C.byref(lr), C.byref(tc), C.byref(ts)))
    ctx.L.fopis_first_solve_cycles.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    ctx._ok(ctx.L.fopis_first_solve_cycles(ctx.h, C.byref(fc)))
    ctx.L.fopis_solver_start_residuals.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    r0 = (C.c_double * 3)()
    ctx._ok(ctx.L.fopis_solver_start_residuals(ctx.h, r0))
    ctx.L.fopis_solver_passes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    pp = (C.c_int64 * 2)()
    ctx._ok(ctx.L.fopis_solver_passes(ctx.h, pp))
    return {"last_cycles": lc.value, "last_relres": lr.value, "total_cycles": tc.value, "total_solves": ts.value,
            "first_cycles": fc.value, "start_relres_last": r0[0], "start_relres_first": r0[1],
            "start_relres_second": r0[2], "updates": pp[0], "pair_passes": pp[1]}
