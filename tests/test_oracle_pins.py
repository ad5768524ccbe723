"""Pins the oracle's remaining functions against what geometry, counting and
statistics fix, independently of the oracle's own formulas:

* conductor capture attribution (Spacecraft::distributeInnerParticleCharge,
  Spacecraft.cpp:512-613, with Particle::computeCrossPoints,
  particle.cpp:82-204): the charge of a particle that ends inside the
  conductor lands on the four nodes of the face square through which its
  straight path entered the conductor, bilinearly at the entry point, which
  is found here by walking the path cell by cell;
* particle energy (GridParticleTreatment.cpp:104-118) and field energy
  (Field.cpp:194-213) against numpy sums over the same state;
* density (RootGridField.cpp:1129-1150) against a numpy histogram;
* the absorbing faces (grid.hpp:309-395) at 0, n-1 and just beyond;
* the initial load (RootGrid.cpp:57-79, AmbientParticleType.cpp:76-93):
  count, uniform positions and normal velocities of the Debye-unit widths.

Each check is also run on an oracle built with one planted mistake
(refcpu.cpp, REFCPU_MUTANT) and must fail there.

The reference's crossing-point order puts x before z for every x-z crossing
(`mvx <= mvx`, particle.cpp:175). With valid cells on both sides of a face,
at most one crossing point of a path satisfies its face predicate, so the
order changes the attribution only for crossing points within 1e-10 of a
second grid plane: the geometric check below cannot tell that order from
the physical one, and does not try to.
"""
import math

import numpy as np
import pytest
from scipy import stats

from fopis_amd import voxel_obj
from oracle import refcpu
from tests.cases import load_config, oracle_params

PARTICLE_KEYS = ("x", "y", "z", "vx", "vy", "vz")


def _grid(cfg, nx, ppc=1, objects=True, lib_path=None, bfield=None):
    return refcpu.Grid(oracle_params(cfg, nx=nx, ppc=ppc, with_objects=objects, bfield=bfield), lib_path=lib_path)


def _mutant(n):
    return refcpu.mutant_path(n)


# ------------------------------------------------------------ capture geometry

def _entry(old, v, inside):
    """Entry of the straight path old -> old + v into the conductor's cells:
    (axis, plane, point) of its single outside -> inside transition, or None
    when the path enters more than once, leaves it, or passes within 1e-9 of
    a second grid plane. `inside(cell)` takes 0-based cell indices (cell c
    spans [c, c + 1))."""
    new = old + v
    cell = np.floor(old).astype(int)
    events = []
    for a in range(3):
        lo, hi = math.floor(old[a]), math.floor(new[a])
        if lo == hi:
            continue
        if abs(hi - lo) != 1:
            return None
        plane = max(lo, hi)
        events.append(((plane - old[a]) / v[a], a, plane, 1 if hi > lo else -1))
    events.sort()
    if len({round(t, 9) for t, *_ in events}) != len(events):
        return None
    was_in = inside(tuple(cell))
    entry = None
    for t, a, plane, step in events:
        cell[a] += step
        now_in = inside(tuple(cell))
        if now_in and not was_in:
            if entry is not None:
                return None
            entry = (a, plane, old + t * v)
        elif was_in and not now_in:
            return None
        was_in = now_in
    if entry is None or not was_in:
        return None
    a, plane, p = entry
    for b in range(3):
        if b != a and abs(p[b] - round(p[b])) < 1e-9:
            return None
    return entry


def _expected_surface_charge(parts, q, inside, node_index, ncmat):
    """Per cmat node: q times the bilinear weights of each path's entry point
    on the four nodes of its entry face square."""
    out = np.zeros(ncmat)
    for old, v in parts:
        a, plane, p = _entry(old, v, inside)
        b, c = [d for d in range(3) if d != a]
        fb, fc = p[b] - math.floor(p[b]), p[c] - math.floor(p[c])
        for db, wb in ((0, 1 - fb), (1, fb)):
            for dc, wc in ((0, 1 - fc), (1, fc)):
                node = [0, 0, 0]
                node[a] = plane + 1  # glue-relative node index of coordinate x is x + 1
                node[b] = math.floor(p[b]) + 1 + db
                node[c] = math.floor(p[c]) + 1 + dc
                out[node_index[tuple(node)]] += q * wb * wc
    return out


def _capture_case(tmp_path, shape):
    """A conductor on a 17^3-node grid and 1500 seeded paths that enter it
    once within one step (|v| < 0.8 per axis)."""
    nx = 17
    cfg = load_config("c2")
    if shape == "cube":
        boxes = [((-2, -2, -2), (2, 2, 2))]
    else:  # L-shape with a concave edge, and a step on the z faces
        boxes = [((-4, -2, -1), (3, 0, 1)), ((-4, 0, -1), (-2, 4, 1)), ((-1, -2, 1), (2, 0, 2))]
    path = str(tmp_path / f"{shape}.obj")
    voxel_obj.write_boxes(path, boxes)
    cfg["Object"]["Spacecraft"]["file_name"] = path
    probe = _grid(cfg, nx)
    cm = probe.cell_map(0).astype(bool)
    inside = lambda c: all(0 <= c[q] < nx - 1 for q in range(3)) and bool(cm[c[0] + 1, c[1] + 1, c[2] + 1])
    rng = np.random.default_rng(12 if shape == "cube" else 13)
    cells = np.argwhere(cm[1:nx, 1:nx, 1:nx])
    parts = []
    while len(parts) < 1500:
        c = cells[rng.integers(len(cells))]
        new = c + rng.uniform(0.02, 0.98, 3)
        v = rng.uniform(-0.8, 0.8, 3)
        old = new - v
        if np.any(old < 0) or np.any(old > nx - 1):
            continue
        if _entry(old, v, inside) is None:
            continue
        parts.append((old, v))
    return cfg, nx, parts, inside


def _check_capture(g, parts, inside):
    nodes = g.cmat_nodes(0)
    node_index = {tuple(int(t) for t in n): i for i, n in enumerate(nodes)}
    ncmat = len(nodes)
    q = g.species_info(0)["q"]
    new = np.array([o + v for o, v in parts])
    vel = np.array([v for _, v in parts])
    d = {k: np.ascontiguousarray(a) for k, a in zip(PARTICLE_KEYS, list(new.T) + list(vel.T))}
    g.set_particles(0, d)
    g.set_particles(1, {k: np.zeros(0) for k in PARTICLE_KEYS})
    g.set_object_charge(0, np.zeros((g.S, ncmat)))
    g.set_object_charge(2, np.zeros(g.S))
    g.call("rc_deposit_only")
    got = g.object_charge(0)[0]
    want = _expected_surface_charge(parts, q, inside, node_index, ncmat)
    assert np.abs(got - want).max() <= 1e-12 * abs(q) * len(parts), np.abs(got - want).max() / abs(q)
    # every path ends inside: all captured, each counted once in the current, no charge left in rho
    assert np.isclose(g.object_charge(2)[0], q * len(parts), rtol=1e-13)
    assert g.get_field(refcpu.RHO, 1).sum() == 0.0


@pytest.mark.parametrize("shape", ["cube", "lshape"])
def test_capture_attribution_is_the_geometric_entry_face(tmp_path, shape):
    cfg, nx, parts, inside = _capture_case(tmp_path, shape)
    _check_capture(_grid(cfg, nx), parts, inside)


def test_capture_check_fails_on_swapped_face_weights(tmp_path):
    cfg, nx, parts, inside = _capture_case(tmp_path, "cube")
    with pytest.raises(AssertionError):
        _check_capture(_grid(cfg, nx, lib_path=_mutant(1)), parts, inside)


# ------------------------------------------------------------ energies

def _state_for_energy(lib_path=None):
    g = _grid(load_config("c1"), 17, ppc=4, objects=False, lib_path=lib_path)
    g.call("rc_load_initial_particles")
    g.call("rc_initialize")
    # one step so that E, phi and the velocities are not the initial ones
    g.call("rc_update_velocity")
    g.call("rc_update_position")
    g.call("rc_update_rho")
    g.call("rc_solve_poisson")
    g.call("rc_update_efield")
    return g


def _check_particle_energy(g):
    want = 0.0
    for s in range(g.S):
        p = g.get_particles(s)
        ok = p["valid"] != 0
        info = g.species_info(s)
        v2 = p["vx"][ok] ** 2 + p["vy"][ok] ** 2 + p["vz"][ok] ** 2
        want += 0.5 * info["mass"] * info["size"] * math.fsum(v2)
    got = g.L.rc_particle_energy(g.h)
    assert abs(got - want) <= 1e-12 * want, (got, want)


def _check_field_energy(g):
    # 1/2 eps0 sum of E^2 over every edge joining two neighbouring nodes of the
    # physical node box (glue excluded), with E on an edge the potential
    # difference of its nodes (dx = 1): computed from phi alone
    phi = g.get_field(refcpu.PHI)[1:-1, 1:-1, 1:-1]
    s = sum(math.fsum((np.diff(phi, axis=a) ** 2).ravel()) for a in range(3))
    want = 0.5 * g.eps0() * s
    got = g.L.rc_efield_energy(g.h)
    assert want > 0 and abs(got - want) <= 1e-11 * want, (got, want)


def test_particle_and_field_energy_are_numpy_sums():
    g = _state_for_energy()
    assert sum((g.get_particles(s)["valid"] == 0).sum() for s in range(g.S)) > 0  # some invalid ones skipped
    _check_particle_energy(g)
    _check_field_energy(g)


def test_energy_checks_fail_on_dropped_terms():
    with pytest.raises(AssertionError):
        _check_particle_energy(_state_for_energy(_mutant(2)))
    with pytest.raises(AssertionError):
        _check_field_energy(_state_for_energy(_mutant(3)))


# ------------------------------------------------------------ density

def _check_density(g):
    cx, cy, cz = g.nx + 2, g.ny + 2, g.nz + 2
    for s in range(g.S):
        p = g.get_particles(s)
        ok = p["valid"] != 0
        want = np.zeros((cx - 1, cy - 1, cz - 1))
        # reference cell of a particle: glue-relative floor(x) + 1 per axis
        idx = [np.floor(p[k][ok]).astype(int) + 1 for k in ("x", "y", "z")]
        np.add.at(want, tuple(idx), g.species_info(s)["size"])
        got = g.get_field(refcpu.DENSITY, s)
        np.testing.assert_array_equal(got, want)


def _density_state(lib_path=None):
    g = _grid(load_config("c1"), 17, ppc=4, objects=False, lib_path=lib_path)
    g.call("rc_load_initial_particles")
    # particles on cell faces, on the upper domain face and on a corner node
    d = g.get_particles(0)
    for k, vals in (("x", [3.0, 16.0, 0.0]), ("y", [5.0, 2.5, 0.0]), ("z", [0.5, 16.0, 16.0])):
        d[k] = np.concatenate([d[k], vals])
    for k in ("vx", "vy", "vz"):
        d[k] = np.concatenate([d[k], [0.0, 0.0, 0.0]])
    d.pop("valid", None)
    g.set_particles(0, d)
    g.call("rc_update_density")
    return g


def test_density_is_a_histogram():
    _check_density(_density_state())


def test_density_check_fails_on_shifted_cells():
    with pytest.raises(AssertionError):
        _check_density(_density_state(_mutant(4)))


# ------------------------------------------------------------ absorbing faces

def _check_absorb(lib_path=None):
    g = _grid(load_config("c1"), 17, ppc=1, objects=False, lib_path=lib_path)
    n1 = g.nx - 1.0
    # after x += v: inside iff 0 <= coordinate <= n - 1 on every axis (Dirichlet faces absorb)
    ends = np.array([-0.5, -2.0 ** -40, 0.0, 0.5, n1 - 2.0 ** -40, n1, np.nextafter(n1, 20.0), n1 + 2.0 ** -40])
    cases = [(e, 4.0, 4.0) for e in ends] + [(4.0, e, 4.0) for e in ends] + [(4.0, 4.0, e) for e in ends]
    end = np.array(cases)
    v = np.full_like(end, 0.25)
    d = {k: np.ascontiguousarray(a) for k, a in zip(PARTICLE_KEYS, list((end - v).T) + list(v.T))}
    g.set_particles(0, d)
    g.set_particles(1, {k: np.zeros(0) for k in PARTICLE_KEYS})
    g.call("rc_update_position")
    p = g.get_particles(0)
    got = p["valid"] != 0
    # (end - v) + v reproduces every end point exactly (v = 1/4, ends chosen for it)
    np.testing.assert_array_equal(np.stack([p["x"], p["y"], p["z"]], 1), end)
    want = np.all((end >= 0.0) & (end <= n1), axis=1)
    np.testing.assert_array_equal(got, want)


def test_absorbing_faces_brute_force():
    _check_absorb()


def test_absorb_check_fails_on_closed_upper_face():
    with pytest.raises(AssertionError):
        _check_absorb(_mutant(5))


# ------------------------------------------------------------ initial load

def _check_load(lib_path=None):
    nx, ppc = 17, 16
    g = _grid(load_config("c1"), nx, ppc=ppc, objects=False, lib_path=lib_path)
    g.call("rc_load_initial_particles")
    # Debye units: dx = lambda_D and w_pe dt = 0.1 give sigma_e = 0.1 dx/dt; T_i = T_e and
    # m_i = 100 m_e give sigma_i = 0.01 dx/dt (the widths sqrt(T/m), not the injection's sqrt(2T/m))
    for s, sigma in ((0, 0.1), (1, 0.01)):
        p = g.get_particles(s)
        n = len(p["x"])
        assert n == (nx - 1) ** 3 * ppc
        for k in ("x", "y", "z"):
            assert p[k].min() >= 0.0 and p[k].max() < nx - 1
            assert stats.kstest(p[k] / (nx - 1), "uniform").pvalue > 1e-4
        for k in ("vx", "vy", "vz"):
            v = p[k]
            assert abs(v.mean()) < 5 * sigma / math.sqrt(n)
            assert abs(v.std() / sigma - 1) < 5 / math.sqrt(2 * n) + 2e-3
            assert stats.kstest(v / sigma, "norm").pvalue > 1e-4


def test_initial_load_law():
    _check_load()


def test_load_check_fails_on_the_injection_width():
    with pytest.raises(AssertionError):
        _check_load(_mutant(6))
