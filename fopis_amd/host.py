"""ctypes binding of the host facade (include/fopis_host.h): the reference
JSON input, the mt19937 particle streams, spacecraft .obj geometry and the
host-side injection/emission around the device step."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as F


class HostOptions(C.Structure):
    _fields_ = [
        ("load_particles", C.c_int32),
        ("device_loader", C.c_int32),
        ("device_seed", C.c_uint64),
        ("ppc_override", C.c_int32),
        ("nx_override", C.c_int32),
        ("build_cmatrix", C.c_int32),
        ("initialize", C.c_int32),
        ("host_threads", C.c_int32),
        ("cmatrix_cache", C.c_char * 512),
        ("no_device", C.c_int32),
        ("device_injection", C.c_int32),
        ("resume_dir", C.c_char * 512),
    ]


_bound = False


def _bind(L):
    global _bound
    if _bound:
        return L
    vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
    L.fopis_host_last_error.restype = C.c_char_p
    L.fopis_host_open.argtypes = [C.c_char_p, C.POINTER(HostOptions), C.POINTER(vp)]
    L.fopis_host_close.argtypes = [vp]
    L.fopis_host_ctx.argtypes = [vp]
    L.fopis_host_ctx.restype = vp
    L.fopis_host_desc.argtypes = [vp, C.POINTER(F.Desc)]
    L.fopis_host_main_loop_es.argtypes = [vp]
    L.fopis_host_timestep.argtypes = [vp]
    L.fopis_host_reload.argtypes = [vp, C.c_uint64]
    L.fopis_host_save_resume.argtypes = [vp, C.c_char_p]
    L.fopis_host_load_resume.argtypes = [vp, C.c_char_p]
    L.fopis_host_generate_ms.argtypes = [vp]
    L.fopis_host_generate_ms.restype = C.c_double
    L.fopis_host_generate_injection.argtypes = [vp, C.c_int, i64] + [dp] * 6
    L.fopis_host_generate_injection.restype = i64
    L.fopis_host_generate_emission.argtypes = [vp, C.c_int, i64] + [dp] * 8
    L.fopis_host_generate_emission.restype = i64
    L.fopis_host_generate_initial.argtypes = [vp, C.c_int, i64] + [dp] * 6
    L.fopis_host_generate_initial.restype = i64
    L.fopis_host_flux.argtypes = [vp, C.c_int, dp]
    L.fopis_host_object.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.c_int32]
    _bound = True
    return L


class _Borrowed(F.Context):
    """A Context view of the facade's device context (not owned)."""

    def __init__(self, handle, desc):
        self.L = F.load()
        self.desc = desc
        self.h = C.c_void_p(handle)
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz
        self.S = desc.nspecies
        self._ncmat = {}

    def close(self):
        self.h = None


class Host:
    """fopis::GpuRootGrid opened on a reference-schema JSON input."""

    def __init__(self, json_path: str, load_particles=True, device_loader=False, device_seed=12345,
                 ppc_override=0, nx_override=0, build_cmatrix=True, initialize=True, cmatrix_cache="",
                 no_device=False, device_injection=False, resume_dir=""):
        self.L = _bind(F.load())
        o = HostOptions()
        o.load_particles = int(load_particles)
        o.device_loader = int(device_loader)
        o.device_seed = device_seed
        o.ppc_override = ppc_override
        o.nx_override = nx_override
        o.build_cmatrix = int(build_cmatrix)
        o.initialize = int(initialize)
        o.cmatrix_cache = cmatrix_cache.encode()
        o.no_device = int(no_device)
        o.device_injection = int(device_injection)
        o.resume_dir = resume_dir.encode()
        h = C.c_void_p()
        rc = self.L.fopis_host_open(json_path.encode(), C.byref(o), C.byref(h))
        if rc != 0:
            raise F.FopisError(rc, self.L.fopis_host_last_error().decode())
        self.h = h
        self.desc = F.Desc()
        self.L.fopis_host_desc(self.h, C.byref(self.desc))
        self.S = self.desc.nspecies
        self.ctx = None
        if not no_device:
            self.ctx = _Borrowed(self.L.fopis_host_ctx(self.h), self.desc)
            for q in range(8):
                try:
                    n = self.object(q)["ncmat"]
                except F.FopisError:
                    break
                self.ctx._ncmat[q] = n

    def close(self):
        if getattr(self, "h", None):
            if self.ctx is not None:
                self.ctx.close()
            self.L.fopis_host_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, rc):
        if rc < 0:
            raise F.FopisError(rc, self.L.fopis_host_last_error().decode())
        return rc

    def main_loop_es(self):
        self._ok(self.L.fopis_host_main_loop_es(self.h))

    def timestep(self) -> int:
        return self.L.fopis_host_timestep(self.h)

    def save_resume(self, directory: str) -> None:
        """RootGrid::saveResumeData + Environment::saveInfo into directory (HDF5)."""
        self._ok(self.L.fopis_host_save_resume(self.h, directory.encode()))

    def load_resume(self, directory: str) -> None:
        """Environment::loadInfo + RootGrid::loadResumeData + initializeForLoad from directory."""
        self._ok(self.L.fopis_host_load_resume(self.h, directory.encode()))

    def reload(self, seed: int) -> None:
        """Redraw the initial particle load on the device with another seed."""
        self._ok(self.L.fopis_host_reload(self.h, seed))

    def generate_ms(self) -> float:
        return self.L.fopis_host_generate_ms(self.h)

    def _gen(self, fn, s):
        n = self._ok(fn(self.h, s, 0, *([None] * 6)))
        return n

    def generate_injection(self, s: int) -> dict:
        """One step of boundary injection for species s (advances the streams)."""
        # the count is only known after drawing, so draw into a generous buffer
        cap = 1 << 22
        arrs = [np.empty(cap) for _ in range(6)]
        n = self._ok(self.L.fopis_host_generate_injection(self.h, s, cap, *[F._dp(a) for a in arrs]))
        if n > cap:
            raise RuntimeError("injection buffer too small")
        return dict(zip(("x", "y", "z", "vx", "vy", "vz"), [a[:n] for a in arrs]))

    def generate_emission(self, s: int):
        """One step of photoemission (every emitting object and species; advances
        the streams): species s's particles, object 0's charge change for s and
        its current change."""
        cap = 1 << 22
        arrs = [np.empty(cap) for _ in range(6)]
        try:
            nc = max(1, self.object(0)["ncmat"])
        except F.FopisError:
            nc = 1
        dq = np.zeros(nc)
        dcur = np.zeros(1)
        n = self._ok(self.L.fopis_host_generate_emission(self.h, s, cap, *[F._dp(a) for a in arrs], F._dp(dq),
                                                         F._dp(dcur)))
        if n > cap:
            raise RuntimeError("emission buffer too small")
        return dict(zip(("x", "y", "z", "vx", "vy", "vz"), [a[:n] for a in arrs])), dq, float(dcur[0])

    def generate_initial(self, s: int) -> dict:
        d = self.desc
        cap = (d.nx - 1) * (d.ny - 1) * (d.nz - 1) * 1024
        cap = min(cap, 1 << 26)
        arrs = [np.empty(cap) for _ in range(6)]
        n = self._ok(self.L.fopis_host_generate_initial(self.h, s, cap, *[F._dp(a) for a in arrs]))
        if n > cap:
            raise RuntimeError("initial-load buffer too small")
        return dict(zip(("x", "y", "z", "vx", "vy", "vz"), [a[:n] for a in arrs]))

    def flux(self, s: int) -> np.ndarray:
        out = np.zeros(6)
        self._ok(self.L.fopis_host_flux(self.h, s, F._dp(out)))
        return out

    def object(self, o: int) -> dict:
        n = C.c_int32()
        self._ok(self.L.fopis_host_object(self.h, o, None, C.byref(n), None, 0))
        d = self.desc
        mask = np.zeros((d.nx + 1, d.ny + 1, d.nz + 1), np.uint8)
        ijk = np.zeros((max(n.value, 1), 3), np.int32)
        self._ok(self.L.fopis_host_object(self.h, o, mask.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n),
                                          ijk.ctypes.data_as(C.POINTER(C.c_int32)), n.value))
        return {"ncmat": n.value, "mask": mask, "cmat_ijk": ijk[: n.value]}
