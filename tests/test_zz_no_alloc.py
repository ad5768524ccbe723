"""fopis_energy_noalloc (the allocation-free variant of fopis_energy) against
fopis_energy through the C ABI: bit-identical results on two calls in a row,
the second on a smaller particle set, and both against the FP64 sums of the
particles and the potential the test itself holds (Grid::getParticleEnergy,
GridParticleTreatment.cpp:104-118; Field::getEfieldEnergy, Field.cpp:194-213)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = pytest.importorskip("fopis_amd.lib")


def _noalloc(ctx):
    L = ctx.L
    L.fopis_energy_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    out = np.zeros(2)
    rc = L.fopis_energy_noalloc(ctx.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, L.fopis_last_error().decode()
    return out


def _expected(parts, species, phi):
    kin = 0.0
    for p, (_, mass, size) in zip(parts, species):
        kin += 0.5 * mass * size * float(np.sum(p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2))
    inner = phi[1:-1, 1:-1, 1:-1]
    ef = (np.sum(np.diff(inner, axis=0) ** 2) + np.sum(np.diff(inner, axis=1) ** 2)
          + np.sum(np.diff(inner, axis=2) ** 2))
    return kin, 0.5 * float(ef)


@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_energy_noalloc_matches_energy(precision):
    nx = 33
    species = [(-1.0, 1.0, 3.0), (1.0, 100.0, 2.0)]  # (charge, mass, size)
    ctx = F.Context(F.make_desc(nx, nx, nx, species, 1.0, precision=precision))
    rng = np.random.default_rng(11)
    try:
        for n in (200_000, 30_000):  # the second call runs on a smaller input
            parts = []
            for s in range(2):
                p = {k: rng.uniform(0.0, nx - 1.0, n) for k in ("x", "y", "z")}
                p.update({k: rng.normal(0.0, 0.1 / (1 + 9 * s), n) for k in ("vx", "vy", "vz")})
                if precision == F.FP32:
                    p = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
                ctx.set_particles(s, p)
                parts.append(p)
            phi = np.zeros(ctx.field_shape(F.PHI))
            phi[1:-1, 1:-1, 1:-1] = rng.normal(size=(nx,) * 3)
            ctx.set_field(F.PHI, phi)
            ref = ctx.energy()
            got = _noalloc(ctx)
            assert got[0] == ref[0] and got[1] == ref[1], (got, ref)
            kin, ef = _expected(parts, species, phi)
            assert abs(got[0] - kin) <= 1e-12 * kin, (got[0], kin)
            assert abs(got[1] - ef) <= 1e-12 * ef, (got[1], ef)
    finally:
        ctx.close()
