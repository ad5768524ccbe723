"""The shipped configurations at their full sizes on the GPU, through the host
facade (the reference's JSON input, RootGrid::mainLoopES), checked step by
step against invariants that hold whatever the particle order or the seed:

* particle ledger: the live count changes by exactly the staged particles
  that joined, minus those absorbed at a domain face, minus those captured by
  the conductor (RootGrid.cpp:82-143, grid.hpp:309-395,
  Spacecraft.cpp:512-592);
* density (RootGridField.cpp:1129-1150): its sum is the superparticle size
  times the live count plus the particles the conductor captured in this
  step, exactly (updateDensity runs before updateRho captures them,
  RootGrid.cpp:115-121);
* charge (RootGridField.cpp:1153-1214): each species' deposit sums to its
  superparticle charge times its live count (FP64 to 1e-12, FP32 to 2e-7
  relative), and rho[0] exceeds the species sum by exactly the conductor's
  total charge, which the capacitance redistribution puts back on the
  surface nodes (Spacecraft.cpp:708-758);
* the conductor's charge changes by the step's current (captured minus
  emitted superparticle charge, Spacecraft.cpp:615-626), less the charge of
  captured particles whose entry face was not found (the reference prints
  "Surface cannot detect" for them and keeps them out of the surface charge,
  Spacecraft.cpp:573-589; a handful per million captures at most);
* every Poisson solve reaches ||b - A phi||_2 / ||b||_2 <= 1e-8.

Plus the device loader's law at the 256^3 size, and the learned-guess start
on the 128^3 grid against the warm start.
"""
import math
import os

import numpy as np
import pytest
from scipy import stats

from tests.cases import CONFIGS

pytestmark = pytest.mark.gpu

F = pytest.importorskip("fopis_amd.lib")
from fopis_amd.host import Host  # noqa: E402


def _open(name, seed=20251202, **kw):
    return Host(os.path.join(CONFIGS, name + ".json"), device_loader=True, device_seed=seed, build_cmatrix=True,
                initialize=True, device_injection=True, **kw)


def _state(h):
    ctx = h.ctx
    out = {"n": [ctx.num_particles(s) for s in range(h.S)], "ledger": [ctx.ledger(s) for s in range(h.S)]}
    out["objects"] = [ctx.object_state(o) for o in sorted(ctx._ncmat) if ctx._ncmat[o] > 0]
    return out


def _check_step(h, before, after, fp32):
    ctx = h.ctx
    st = F.solver_stats(ctx)
    assert 0 < st["last_relres"] <= 1e-8, st
    for s in range(h.S):
        d = {k: after["ledger"][s][k] - before["ledger"][s][k] for k in ("joined", "absorbed", "captured")}
        assert after["n"][s] == before["n"][s] + d["joined"] - d["absorbed"] - d["captured"], (s, before, after)
        sd = h.desc.species[s]
        dens = ctx.get_field(F.DENSITY, s)
        assert dens.sum() == (after["n"][s] + d["captured"]) * sd.size
        rho = ctx.get_field(F.RHO, s + 1)
        want = sd.charge * sd.size * after["n"][s]
        tol = 2e-7 if fp32 else 1e-12
        assert abs(rho.sum() - want) <= tol * abs(want) + 1e-300, (s, rho.sum(), want)
    rho0 = ctx.get_field(F.RHO, 0).sum()
    species_sum = sum(ctx.get_field(F.RHO, s + 1).sum() for s in range(h.S))
    q_obj = sum(o["total_charge"] for o in after["objects"])
    scale = max(abs(species_sum), abs(q_obj), sum(abs(h.desc.species[s].charge * h.desc.species[s].size)
                                                  * after["n"][s] for s in range(h.S)) * 1e-12)
    assert abs(rho0 - species_sum - q_obj) <= 1e-9 * scale, (rho0, species_sum, q_obj)
    for ob0, ob1 in zip(before["objects"], after["objects"]):
        dq = ob1["total_charge"] - ob0["total_charge"]
        cur = float(np.sum(ob1["current"]))
        lost = ob1["unattributed"]
        q_lost = sum(float(lost[s]) * h.desc.species[s].charge * h.desc.species[s].size for s in range(h.S))
        assert abs(dq - (cur - q_lost)) <= 1e-9 * max(abs(cur), abs(ob1["total_charge"]), 1.0), (dq, cur, lost)
        captured = sum(after["ledger"][s]["captured"] - before["ledger"][s]["captured"] for s in range(h.S))
        assert int(lost.sum()) <= max(2, captured // 1000), (lost, captured)


def _run(name, steps, fp32):
    h = _open(name)
    try:
        before = _state(h)
        assert sum(before["n"]) > 0
        for _ in range(steps):
            h.main_loop_es()
            after = _state(h)
            _check_step(h, before, after, fp32)
            before = after
        return after
    finally:
        h.close()


def test_c2_full_size_steps():
    """128^3 cells, 64 ppc x 2 species (FP64), floating 16^3 cube, x-face injection."""
    end = _run("c2", 3, fp32=False)
    assert end["n"][0] > 1.3e8 and end["n"][1] > 1.3e8
    led = end["ledger"]
    assert led[1]["joined"] > 0 and led[0]["absorbed"] > 0 and led[0]["captured"] > 0


def test_c3_full_size_steps():
    """256^3 cells, 64 ppc x 2 species (2.15e9 FP32 particles), 32^3 cube, B0 = 0.2 z."""
    end = _run("c3", 2, fp32=True)
    assert sum(end["n"]) > 2.1e9
    assert end["ledger"][0]["captured"] > 0


def test_c5_full_size_steps():
    """256^3 cells, L-shaped spacecraft, photoelectrons from its lit face."""
    end = _run("c5", 2, fp32=True)
    assert len(end["n"]) == 3
    assert end["ledger"][2]["joined"] > 0  # photoelectrons joined the plasma
    assert end["objects"][0]["current"][2] != 0.0


def _species_law(name):
    """sigma and drift of each species in dx/dt from the oracle's normalisation
    (itself pinned to the Debye-unit widths, tests/test_oracle.py)."""
    from oracle import refcpu
    from tests.cases import load_config, oracle_params

    g = refcpu.Grid(oracle_params(load_config(name), nx=9, ppc=1, with_objects=False))
    return [(g.species_info(s)["deviation"], tuple(g.species_info(s)["drift"])) for s in range(g.S)]


def test_device_loader_law():
    """fopis_load_maxwellian (the law of RootGrid.cpp:57-79 and
    AmbientParticleType.cpp:76-93: positions uniform over the box, velocities
    N(drift, sigma) per component; Spacecraft::removeInnerParticle drops
    those born inside a conductor): KS tests on every component, moments,
    the count, and no particle in a conductor cell. 64^3 cells, 16 ppc."""
    nx, ppc = 65, 16
    ctx = F.Context(F.make_desc(nx, nx, nx, [(-1.0, 1.0, 1.0), (1.0, 100.0, 1.0)], 1.0))
    mask = np.zeros((nx + 1,) * 3, np.uint8)
    mask[28:38, 30:34, 25:40] = 1  # a 10 x 4 x 15-cell box (glue-relative cells)
    try:
        ctx.set_object(0, mask, np.array([[30, 31, 27]], np.int32))
        lo, hi = (0.0, 0.0, 0.0), (nx - 1.0,) * 3
        for s, (sigma, drift) in enumerate(((0.1, (0.0, 0.0, 0.0)), (0.01, (0.05, -0.02, 0.0)))):
            n = (nx - 1) ** 3 * ppc
            ctx.load_maxwellian(s, n, 1234 + s, lo, hi, drift, sigma)
            p = ctx.get_particles(s)
            m = len(p["x"])
            masked = mask[1:-1, 1:-1, 1:-1].sum()
            # dropped ~ Binomial(n, masked / cells)
            pr = masked / (nx - 1) ** 3
            assert abs((n - m) - n * pr) <= 6 * math.sqrt(n * pr), (n - m, n * pr)
            ci = tuple(np.floor(p[k]).astype(int) + 1 for k in ("x", "y", "z"))
            assert not mask[ci].any()
            for k in ("x", "y", "z"):
                assert p[k].min() >= 0.0 and p[k].max() < nx - 1
            for q, k in enumerate(("vx", "vy", "vz")):
                v = p[k]
                assert abs(v.mean() - drift[q]) < 5 * sigma / math.sqrt(m), (s, k, v.mean())
                assert abs(v.std() / sigma - 1) < 5 / math.sqrt(2 * m), (s, k, v.std())
                assert stats.kstest((v - drift[q]) / sigma, "norm").pvalue > 1e-4
            # positions: KS on x over the particles whose (y, z) column misses the conductor
            free = ~((p["y"] >= 29) & (p["y"] < 33) & (p["z"] >= 24) & (p["z"] < 39))
            for k in ("x", "y", "z"):
                sel = free if k == "x" else ~((p["x"] >= 27) & (p["x"] < 37))
                assert stats.kstest(p[k][sel] / (nx - 1), "uniform").pvalue > 1e-4, k
    finally:
        ctx.close()


def test_device_loader_at_256():
    """The C3 load (2.15e9 FP32 particles, 32^3 cube): count, phase-space
    moments over every particle (fopis_moments), and no density in a
    conductor cell."""
    law = _species_law("c3")
    h = _open("c3")
    try:
        ctx = h.ctx
        mask = h.object(0)["mask"].astype(bool)
        inside = int(np.count_nonzero(mask))
        for s in range(h.S):
            n = ctx.num_particles(s)
            drawn = 256 ** 3 * 64
            pr = inside / 256 ** 3
            assert abs((drawn - n) - drawn * pr) <= 6 * math.sqrt(drawn * pr), (drawn - n, drawn * pr)
            assert not ctx.get_field(F.DENSITY, s)[mask].any()
            m = ctx.moments(s)
            sigma, drift = law[s]
            for q in range(3):
                # positions: uniform over [0, 256) outside the centred cube, whose
                # removal leaves the mean at the centre
                assert abs(m["mean_x"][q] - 128.0) < 5 * (256 / math.sqrt(12)) / math.sqrt(n)
                assert abs(m["mean_v"][q] - drift[q]) < 5 * sigma / math.sqrt(n)
                assert abs(m["var_v"][q] / sigma ** 2 - 1) < 5 * math.sqrt(2 / n) + 1e-6
    finally:
        h.close()


def test_c4_learned_correction_against_the_warm_start():
    """The learned-guess workload (128^3, the C2 plasma around its 16^3 cube,
    FP32 particles) with the shipped network trained on warm-start pairs
    (fopis_amd/weights/guess_c4_corr.npz, guess mode 2): the first solve of
    each step starts from a relative residual at least ten times below the
    warm start's (the reference warm-starts from the previous phi,
    RootGridField.cpp:41-103), needs no more V-cycles, and ends at 1e-8. The
    wall-clock comparison (inference included) is the bench's learned leg."""
    import json

    from fopis_amd.guess import Guess

    w = os.path.join(os.path.dirname(CONFIGS), "weights", "guess_c4_corr.npz")
    z = np.load(w, allow_pickle=False)
    assert json.loads(str(z["recipe"]))["recipe"] == "correction"
    with open(os.path.join(CONFIGS, "c2.json")) as f:
        cfg = json.load(f)
    cfg["Build"]["precision"] = "fp32"
    for o in cfg["Object"].values():
        o["file_name"] = os.path.join(CONFIGS, o["file_name"])
    import tempfile

    fd, path = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(cfg, f)
    h = Host(path, device_loader=True, device_seed=4242, build_cmatrix=True, initialize=True, device_injection=True)
    os.unlink(path)
    try:
        ctx = h.ctx
        g = Guess(ctx, seed=1)
        g.set_params(z["params"], float(z["out_scale"]))
        for _ in range(2):
            h.main_loop_es()
        res = {}
        for mode in (0, 2):
            g.set_mode(mode)
            ctx.reset_timers()
            for _ in range(4):
                h.main_loop_es()
                assert F.solver_stats(ctx)["last_relres"] <= 1e-8
            st = F.solver_stats(ctx)
            res[mode] = (st["start_relres_first"], st["first_cycles"] / 4, ctx.phase_time("poisson_first")[0])
        print(f"[c4] warm start: relres {res[0][0]:.3g}, {res[0][1]} cycles, {res[0][2]:.3f} ms; "
              f"network: relres {res[2][0]:.3g}, {res[2][1]} cycles, {res[2][2]:.3f} ms (inference included)")
        assert res[2][0] < 0.1 * res[0][0]
        assert res[2][1] <= res[0][1]
    finally:
        h.close()
