"""Parity of the HIP path (through the C ABI) with the oracle on identical
seeded inputs.

Bars: bit-exact for FP64 push/move and cell keys, E from identical phi and
the conductor mask; relative 1e-12 (FP64) for charge density, exact for the
density counts; Poisson within the stated residual of the exact discrete
solution."""
import numpy as np
import pytest

from oracle import refcpu
from tests.cases import canon, device_desc, dst_poisson, load_config, oracle_params

pytestmark = pytest.mark.gpu

F = pytest.importorskip("fopis_amd.lib")


def _oracle(nx=33, ppc=4, b=(0.0, 0.0, 0.0), cfg="c1", with_objects=False):
    g = refcpu.Grid(oracle_params(load_config(cfg), nx=nx, ppc=ppc, with_objects=with_objects, bfield=b))
    g.call("rc_load_initial_particles")
    return g


def _random_eref(g, seed=1, scale=0.05):
    rng = np.random.default_rng(seed)
    shp = g.field_shape(refcpu.EXREF)
    return [rng.uniform(-scale, scale, shp) for _ in range(3)]


def _ctx(g, precision=F.FP64, **kw):
    return F.Context(device_desc(g, precision=precision, **kw))


def _upload(ctx, g):
    for s in range(g.S):
        ctx.set_particles(s, g.get_particles(s))


def tile_order(d, nx):
    """Device order after set_particles: stable by 8^3-cell tile of the position."""
    N = nx - 1
    T = -(-N // 8)
    c = [np.minimum(np.floor(d[k]).astype(np.int64), N - 1) // 8 for k in ("x", "y", "z")]
    return np.argsort((c[0] * T + c[1]) * T + c[2], kind="stable")


def _stack(d, order=None):
    a = np.stack([np.asarray(d[k]) for k in ("x", "y", "z", "vx", "vy", "vz")], axis=1)
    return a if order is None else a[order]


def test_device_info():
    info = F.device_info()
    assert "gfx950" in info, info


def test_roundtrip_particles():
    g = _oracle()
    ctx = _ctx(g)
    _upload(ctx, g)
    for s in range(g.S):
        assert ctx.num_particles(s) == g.num_particles(s)
        np.testing.assert_array_equal(canon(ctx.get_particles(s)), canon(g.get_particles(s)))


@pytest.mark.parametrize("bfield", [(0.0, 0.0, 0.0), (300.0, -150.0, 2028.6)])
def test_push_velocity_bitexact_fp64(bfield):
    g = _oracle(b=bfield)
    ctx = _ctx(g)
    _upload(ctx, g)
    E = _random_eref(g)
    for q, w in enumerate((refcpu.EXREF, refcpu.EYREF, refcpu.EZREF)):
        g.set_field(w, E[q])
        ctx.set_field(F.EXREF + q, E[q])
    order = {s: tile_order(g.get_particles(s), g.nx) for s in range(g.S)}
    g.call("rc_update_velocity")
    ctx.update_particle_velocity_es()
    for s in range(g.S):
        np.testing.assert_array_equal(_stack(ctx.get_particles(s)), _stack(g.get_particles(s), order[s]))


def test_push_position_bitexact_fp64():
    """Velocity then position, twice, through the split C ABI without
    re-binning: device order is the initial tile order. The absorbed
    particles stay in their slots on the device until the next re-bin; the
    live count, the copy-out, the energy (Grid::getValidParticleNumber,
    Grid::getParticleEnergy, GridParticleTreatment.cpp:104-128) and the
    ledger all leave them out, as the reference's isValid flag does."""
    g = _oracle(b=(0.0, 0.0, 2028.6))
    ctx = _ctx(g)
    _upload(ctx, g)
    order = {s: tile_order(g.get_particles(s), g.nx) for s in range(g.S)}
    E = _random_eref(g, seed=3, scale=0.3)
    for q, w in enumerate((refcpu.EXREF, refcpu.EYREF, refcpu.EZREF)):
        g.set_field(w, E[q])
        ctx.set_field(F.EXREF + q, E[q])
    for _ in range(2):
        g.call("rc_update_velocity")
        g.call("rc_update_position")
        ctx.update_particle_velocity_es()
        ctx.update_particle_position_es()
    for s in range(g.S):
        ref = g.get_particles(s)
        a = _stack(ctx.get_particles(s))
        b = _stack(ref, order[s])
        valid = ref["valid"][order[s]] != 0
        assert valid.sum() < len(valid)  # some particles left through a face
        assert ctx.num_particles(s) == valid.sum() == len(a)
        assert ctx.ledger(s)["absorbed"] == (~valid).sum()
        np.testing.assert_array_equal(a, b[valid])
    e_dev = ctx.energy()[0]
    e_ref = g.L.rc_particle_energy(g.h)
    assert abs(e_dev - e_ref) <= 1e-12 * e_ref


def test_deposit_density_fp64():
    g = _oracle(ppc=8)
    ctx = _ctx(g)
    _upload(ctx, g)
    g.call("rc_update_density")
    g.call("rc_deposit_only")
    ctx.update_density()
    ctx.update_rho()
    for s in range(g.S):
        np.testing.assert_array_equal(ctx.get_field(F.DENSITY, s), g.get_field(refcpu.DENSITY, s))
    for s in range(g.S + 1):
        a, b = ctx.get_field(F.RHO, s), g.get_field(refcpu.RHO, s)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), s
    # charge conservation: total deposited charge equals sum of q*size
    for s in range(g.S):
        q = g.species_info(s)["q"]
        tot = ctx.get_field(F.RHO, s + 1).sum()
        assert abs(tot - q * g.num_particles(s)) <= 1e-12 * abs(q * g.num_particles(s))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_deposit_is_order_independent(precision):
    """The tile charge windows sum fixed-point weights with integer atomics,
    so rho does not depend on the order the particles are stored (or the
    atomics land) in: the same particles in a shuffled order give rho bit
    for bit."""
    g = _oracle(ppc=8)
    prec = F.FP64 if precision == "fp64" else F.FP32
    a, b = _ctx(g, precision=prec), _ctx(g, precision=prec)
    rng = np.random.default_rng(4)
    for s in range(g.S):
        d = g.get_particles(s)
        perm = rng.permutation(len(d["x"]))
        a.set_particles(s, d)
        b.set_particles(s, {k: np.asarray(v)[perm] for k, v in d.items() if k != "valid"})
    a.update_rho()
    b.update_rho()
    for s in range(g.S + 1):
        np.testing.assert_array_equal(a.get_field(F.RHO, s), b.get_field(F.RHO, s))
    a.close()
    b.close()


def test_fused_step_matches_stage_by_stage():
    """fopis_main_loop_es (one fused kernel per species) against the stage-by-stage entry points."""
    g = _oracle(ppc=6, b=(0.0, 0.0, 2028.6))
    a, b = _ctx(g), _ctx(g)
    _upload(a, g)
    _upload(b, g)
    a.initialize()
    b.initialize()
    # give both contexts the same E before the step (the stage-by-stage path
    # solves with its own cycle batching)
    for q in range(3):
        b.set_field(F.EXREF + q, a.get_field(F.EXREF + q))
    a.main_loop_es()
    b.update_particle_velocity_es()
    b.update_particle_position_es()
    b.update_density()
    b.update_rho()
    b.solve_poisson()
    b.update_efield()
    for s in range(g.S):
        np.testing.assert_array_equal(canon(a.get_particles(s)), canon(b.get_particles(s)))
        np.testing.assert_array_equal(a.get_field(F.DENSITY, s), b.get_field(F.DENSITY, s))
    ra, rb = a.get_field(F.RHO, 0), b.get_field(F.RHO, 0)
    assert np.max(np.abs(ra - rb)) <= 1e-12 * np.max(np.abs(rb))


@pytest.mark.parametrize("mgp", [F.FP64, F.FP32])
def test_poisson_against_exact(mgp):
    g = _oracle(ppc=8)
    g.call("rc_deposit_only")
    rho = g.get_field(refcpu.RHO, 0)
    ctx = _ctx(g, mg_precision=mgp)
    ctx.set_field(F.RHO, rho, 0)
    cyc, rr = ctx.solve_poisson()
    assert rr <= 1e-8 and cyc < 30
    phi = ctx.get_field(F.PHI)
    b = rho[2:-2, 2:-2, 2:-2] / g.eps0()
    exact = dst_poisson(b)
    err = np.linalg.norm(phi[2:-2, 2:-2, 2:-2] - exact) / np.linalg.norm(exact)
    assert err < 1e-6, err
    # the reference's own SOR run to the same criterion lands on the same solution
    g.call("rc_solve_poisson_relres", 1e-10, 20000)
    ref = g.get_field(refcpu.PHI)
    assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) < 1e-6


def test_efield_bitexact_from_same_phi():
    g = _oracle(ppc=2)
    rng = np.random.default_rng(7)
    phi = np.zeros(g.field_shape(refcpu.PHI))
    phi[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(s - 2 for s in phi.shape))
    g.set_field(refcpu.PHI, phi)
    g.call("rc_update_efield")
    ctx = _ctx(g)
    ctx.set_field(F.PHI, phi)
    ctx.update_efield()
    for q, w in enumerate((refcpu.EXREF, refcpu.EYREF, refcpu.EZREF)):
        np.testing.assert_array_equal(ctx.get_field(F.EXREF + q), g.get_field(w))
    for q, w in enumerate((refcpu.EX, refcpu.EY, refcpu.EZ)):
        np.testing.assert_array_equal(ctx.get_field(F.EX + q), g.get_field(w))


def test_sort_by_cell_is_stable_counting_sort():
    g = _oracle(ppc=5)
    ctx = _ctx(g)
    _upload(ctx, g)
    before = {s: ctx.get_particles(s) for s in range(g.S)}
    ctx.sort_by_cell()
    for s in range(g.S):
        d = before[s]
        x, y, z = d["x"], d["y"], d["z"]
        N = g.nx - 1
        ci = np.minimum(np.floor(x).astype(int), N - 1)
        cj = np.minimum(np.floor(y).astype(int), N - 1)
        ck = np.minimum(np.floor(z).astype(int), N - 1)
        T = -(-N // 8)
        tile = ((ci // 8) * T + cj // 8) * T + ck // 8
        key = tile * 512 + ((ci % 8) * 8 + cj % 8) * 8 + ck % 8
        order = np.argsort(key, kind="stable")
        after = ctx.get_particles(s)
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(after[k], d[k][order])
        keys = ctx.cell_keys(s)
        ref_keys = ((np.floor(after["x"]).astype(int) + 1) * (g.ny + 1) + np.floor(after["y"]).astype(int) + 1) * (
            g.nz + 1) + np.floor(after["z"]).astype(int) + 1
        np.testing.assert_array_equal(keys, ref_keys)


def test_multi_step_against_oracle_fp64():
    g = _oracle(ppc=4)
    ctx = _ctx(g)
    _upload(ctx, g)
    g.call("rc_initialize")
    ctx.initialize()
    for _ in range(3):
        g.call("rc_main_loop")
        ctx.main_loop_es()
    for s in range(g.S):
        a, b = canon(ctx.get_particles(s)), canon(g.get_particles(s))
        assert a.shape == b.shape
        assert np.max(np.abs(a - b)) < 1e-9
    ke_gpu, fe_gpu = ctx.energy()
    assert abs(ke_gpu - g.L.rc_particle_energy(g.h)) <= 1e-9 * abs(ke_gpu)
    assert abs(fe_gpu - g.L.rc_efield_energy(g.h)) <= 1e-5 * abs(fe_gpu)


def test_forty_steps_track_the_oracle_fp64():
    """Forty whole steps (no conductor, no injection): the same particles
    leave through the faces on both sides, and the particle state, the
    kinetic and the field energy stay with the oracle's. The two Poisson
    solvers agree to ~1e-10, and the difference grows to ~1e-8 in 40 steps;
    the bars leave a factor of 100."""
    g = _oracle(ppc=4)
    ctx = _ctx(g)
    _upload(ctx, g)
    g.call("rc_initialize")
    ctx.initialize()
    for _ in range(40):
        g.call("rc_main_loop")
        ctx.main_loop_es()
    for s in range(g.S):
        a, b = canon(ctx.get_particles(s)), canon(g.get_particles(s))
        assert a.shape == b.shape
        assert np.max(np.abs(a - b)) < 1e-6
    assert ctx.num_particles(0) < (g.nx - 1) ** 3 * 4 * 0.9  # over a tenth of the electrons left
    ke, fe = ctx.energy()
    assert abs(ke - g.L.rc_particle_energy(g.h)) <= 1e-8 * abs(ke)
    assert abs(fe - g.L.rc_efield_energy(g.h)) <= 1e-6 * abs(fe)
    ctx.close()


def test_far_movers_and_crowded_tile_match_oracle():
    """Particles that jump more than a tile in one step go through the stray
    list and deposit outside their tile's charge window; a crowd converging
    on one cell brings the target tile more arrivals than its binned
    capacity holds (the rest spill to the stray list). Over three full steps the particles follow the
    oracle's (GridParticleTreatmentES.cpp:13-181) to the multi-step test's bar
    of 1e-9 (each side's field comes from its own Poisson solve), and after
    every step the density counts and charge density equal the oracle's
    deposit of the same particles (RootGridField.cpp:1129-1252): counts
    exactly, rho to 1e-12 relative."""
    g = _oracle(ppc=4)
    rng = np.random.default_rng(11)
    for s in range(g.S):
        d = g.get_particles(s)
        n = len(d["x"])
        far = rng.random(n) < 0.1
        ax = rng.integers(0, 3, n)
        speed = rng.uniform(9.0, 12.0, n) * rng.choice([-1.0, 1.0], n)
        for k, key in enumerate(("vx", "vy", "vz")):
            d[key] = np.where(far & (ax == k), speed, d[key])
        pos = np.stack([d["x"], d["y"], d["z"]])
        crowd = ~far & np.all(np.abs(pos - 20.0) < 6.0, axis=0)
        for k, key in enumerate(("vx", "vy", "vz")):
            d[key] = np.where(crowd, 20.5 - pos[k], d[key])
        assert far.sum() > 1000 and crowd.sum() > 1000
        g.set_particles(s, d)
    ctx = _ctx(g)
    _upload(ctx, g)
    g.call("rc_initialize")
    ctx.initialize()
    for _ in range(3):
        g.call("rc_main_loop")
        ctx.main_loop_es()
        h = _oracle(ppc=4)
        for s in range(g.S):
            d = ctx.get_particles(s)
            a, b = canon(d), canon(g.get_particles(s))
            assert a.shape == b.shape
            assert np.max(np.abs(a - b)) < 1e-9
            h.set_particles(s, d)
        h.call("rc_update_density")
        h.call("rc_deposit_only")
        for s in range(g.S):
            np.testing.assert_array_equal(ctx.get_field(F.DENSITY, s), h.get_field(refcpu.DENSITY, s))
        for s in range(g.S + 1):
            a, b = ctx.get_field(F.RHO, s), h.get_field(refcpu.RHO, s)
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(h.get_field(refcpu.RHO, 1))), s


def test_empty_species_steps():
    """A species with no particles (both on the device and in the oracle) runs
    through the fused step: its density and charge stay zero and the other
    species keeps the oracle's particles."""
    g = _oracle(ppc=4)
    empty = {k: np.zeros(0) for k in ("x", "y", "z", "vx", "vy", "vz")}
    g.set_particles(1, empty)
    ctx = _ctx(g)
    _upload(ctx, g)
    assert ctx.num_particles(1) == 0
    g.call("rc_initialize")
    ctx.initialize()
    for _ in range(2):
        g.call("rc_main_loop")
        ctx.main_loop_es()
    assert ctx.num_particles(1) == 0 and g.num_particles(1) == 0
    assert not np.any(ctx.get_field(F.DENSITY, 1)) and not np.any(ctx.get_field(F.RHO, 2))
    a, b = canon(ctx.get_particles(0)), canon(g.get_particles(0))
    assert a.shape == b.shape and np.max(np.abs(a - b)) < 1e-9


def test_particles_on_faces_and_node_lines_deposit_exactly():
    """Particles exactly on the domain faces (x = 0 and x = n - 1 on each
    axis, the largest valid coordinate), on node lines and at cell corners
    deposit and count like the oracle (position.hpp:13-69: the cell index
    floor(x) + 1 is clamped at the upper face)."""
    g = _oracle(ppc=2)
    N = float(g.nx - 1)
    rng = np.random.default_rng(9)
    for s in range(g.S):
        d = {k: np.asarray(v, np.float64).copy() for k, v in g.get_particles(s).items() if k != "valid"}
        m = 3000
        special = rng.choice([0.0, N, 1.0, 8.0, 16.0, N - 1.0], size=(3, m))
        free = rng.uniform(0.0, N, (3, m))
        pick = rng.random((3, m)) < 0.6
        pos = np.where(pick, special, free)
        for k, key in enumerate(("x", "y", "z")):
            d[key][:m] = pos[k]
        g.set_particles(s, d)
    ctx = _ctx(g)
    _upload(ctx, g)
    g.call("rc_update_density")
    g.call("rc_deposit_only")
    ctx.update_density()
    ctx.update_rho()
    for s in range(g.S):
        np.testing.assert_array_equal(ctx.get_field(F.DENSITY, s), g.get_field(refcpu.DENSITY, s))
    for s in range(g.S + 1):
        a, b = ctx.get_field(F.RHO, s), g.get_field(refcpu.RHO, s)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), s


def test_fp32_push_close_to_fp64():
    g = _oracle(ppc=4, b=(0.0, 0.0, 2028.6))
    ctx = _ctx(g, precision=F.FP32)
    _upload(ctx, g)
    E = _random_eref(g, seed=5, scale=0.05)
    for q, w in enumerate((refcpu.EXREF, refcpu.EYREF, refcpu.EZREF)):
        g.set_field(w, E[q])
        ctx.set_field(F.EXREF + q, E[q])
    order = {s: tile_order(g.get_particles(s), g.nx) for s in range(g.S)}
    g.call("rc_update_velocity")
    g.call("rc_update_position")
    ctx.update_particle_velocity_es()
    ctx.update_particle_position_es()
    for s in range(g.S):
        ref = g.get_particles(s)
        a = _stack(ctx.get_particles(s))
        b = _stack(ref, order[s])
        valid = ref["valid"][order[s]] != 0
        assert valid.mean() > 0.99
        # the same particles left through the faces (none ends within FP32 rounding of one)
        assert len(a) == valid.sum()
        b = b[valid]
        # FP32 push against the FP64 oracle: positions < 32 carry at most a few
        # ulps of 2^-19; velocities agree to FP32 relative precision
        assert np.max(np.abs(a[:, :3] - b[:, :3])) < 4 * 2.0 ** -19
        assert np.max(np.abs(a[:, 3:] - b[:, 3:])) < 1e-6


def test_fp32_fused_step_deposit():
    """The FP32 fused kernel (push + move + count + deposit + bin in one launch)
    deposits the charge of the particles it moved: the oracle's CIC deposit of
    the same post-step particles agrees to FP32 weight rounding (the windows
    accumulate in FP64), and the density counts agree exactly."""
    g = _oracle(ppc=6, b=(0.0, 0.0, 2028.6))
    ctx = _ctx(g, precision=F.FP32)
    _upload(ctx, g)
    ctx.initialize()
    ctx.main_loop_es()
    for s in range(g.S):
        d = ctx.get_particles(s)
        keep = ~np.isnan(np.asarray(d["x"]))
        g.set_particles(s, {k: np.asarray(d[k], dtype=np.float64)[keep] for k in ("x", "y", "z", "vx", "vy", "vz")})
    g.call("rc_update_density")
    g.call("rc_deposit_only")
    for s in range(g.S):
        np.testing.assert_array_equal(ctx.get_field(F.DENSITY, s), g.get_field(refcpu.DENSITY, s))
        a, b = ctx.get_field(F.RHO, s + 1), g.get_field(refcpu.RHO, s + 1)
        # each particle's eight FP32 weights carry ~3 roundings of 2^-24
        assert np.max(np.abs(a - b)) <= 2e-6 * np.max(np.abs(b)), s


def test_moments_match_host_reduction():
    g = _oracle(ppc=4)
    for prec in (F.FP64, F.FP32):
        ctx = _ctx(g, precision=prec)
        _upload(ctx, g)
        for s in range(g.S):
            d = ctx.get_particles(s)
            m = ctx.moments(s)
            a = np.stack([np.asarray(d[k], np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")])
            assert m["count"] == a.shape[1]
            np.testing.assert_allclose(m["mean_x"], a[:3].mean(axis=1), rtol=1e-12)
            np.testing.assert_allclose(m["mean_v"], a[3:].mean(axis=1), rtol=1e-9, atol=1e-15)
            np.testing.assert_allclose(m["var_v"], a[3:].var(axis=1), rtol=1e-9)
        ctx.close()


@pytest.mark.parametrize("mg_precision", [F.FP32, F.FP64])
def test_poisson_graph_replays_match_eager_launches(monkeypatch, mg_precision):
    """The V-cycle captured as a graph (kernel nodes only) and replayed over
    several solves, with eager work and synchronisations between them, gives
    the potential of the same cycles launched kernel by kernel, bit for bit
    (every reduction of the solve runs in a fixed order)."""
    g = _oracle(nx=33, ppc=8)
    g.call("rc_deposit_only")
    rho = g.get_field(refcpu.RHO, 0)
    ctxs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FOPIS_MG_GRAPH", mode)
        ctxs[mode] = _ctx(g, mg_precision=mg_precision)
    rng = np.random.default_rng(4)
    out = {m: [] for m in ctxs}
    for _ in range(4):
        scale = rng.uniform(0.5, 2.0)
        noise = rng.normal(size=rho.shape) * 0.05 * np.abs(rho).max()
        for m, ctx in ctxs.items():
            ctx.set_field(F.RHO, rho * scale + noise, 0)
            cyc, rr = ctx.solve_poisson()
            ctx.update_efield()  # eager work between the replays
            ctx.synchronize()
            out[m].append((cyc, rr, ctx.get_field(F.PHI), ctx.get_field(F.EXREF)))
    for a, b in zip(out["1"], out["0"]):
        assert a[0] == b[0] and a[0] > 0 and a[1] == b[1] <= 1e-8
        np.testing.assert_array_equal(a[2], b[2])
        np.testing.assert_array_equal(a[3], b[3])
    for ctx in ctxs.values():
        ctx.close()


@pytest.mark.parametrize("pairs", ["1", "0"])
def test_colour_split_finest_level_matches_one_array(monkeypatch, pairs):
    """An FP32 finest level of 127^3 unknowns (2^20 or more: it forms its
    residual in a pass of its own) is stored by colour, red and black halves
    apart; every kernel on it forms the same sums in the same order as on the
    one-array layout (FOPIS_MG_RB=0), so successive warm-started solves give
    the same potential bit for bit and the same cycles, with single cycles and
    with pairs; the residual norms agree to rounding (the split passes add up
    a row's points in another order)."""
    n = 129
    monkeypatch.setenv("FOPIS_MG_PAIRS", pairs)
    ctxs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FOPIS_MG_RB", mode)
        ctxs[mode] = F.Context(F.make_desc(n, n, n, [(-1.0, 1.0, 1.0)], eps0=1.0, precision=F.FP32,
                                           mg_precision=F.FP32))
    rng = np.random.default_rng(11)
    shape = ctxs["1"].field_shape(F.RHO)
    z, y, x = np.meshgrid(*[np.linspace(0, 1, s) for s in shape], indexing="ij")
    base = np.sin(3 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(np.pi * z)
    out = {m: [] for m in ctxs}
    for _ in range(3):
        rho = base * rng.uniform(0.5, 2.0) + rng.normal(size=shape) * 0.2
        for m, ctx in ctxs.items():
            ctx.set_field(F.RHO, rho, 0)
            cyc, rr = ctx.solve_poisson()  # warm-started from the previous potential
            out[m].append((cyc, rr, ctx.get_field(F.PHI)))
    for a, b in zip(out["1"], out["0"]):
        assert a[0] == b[0] and a[0] > 0 and a[1] <= 1e-8 and b[1] <= 1e-8, (a[:2], b[:2])
        assert abs(a[1] - b[1]) <= 1e-12 * b[1], (a[:2], b[:2])
        np.testing.assert_array_equal(a[2], b[2])
    for ctx in ctxs.values():
        ctx.close()
