"""The C++ host facade (fopis::GpuRootGrid, include/fopis_host.h) against the
oracle on the CPU: the normalised description it hands the device, the
initial load, boundary injection and photoemission drawn from the
reference's mt19937 streams (bit-exact), and the spacecraft geometry read from .obj files
(bit-exact mask and cmat numbering). Also: the C ABI library loads and
exports every declared entry point."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from oracle import refcpu
from tests.cases import CONFIGS, GOLDEN, load_config, oracle_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib_path():
    return os.path.join(ROOT, "fopis_amd", "libfopis_hip.so")


def test_library_exports_declared_entry_points():
    lib = ctypes.CDLL(_lib_path())
    io_path = os.path.join(ROOT, "fopis_amd", "libfopis_io.so")
    for header, L in (("fopis.h", lib), ("fopis_host.h", lib),
                      ("fopis_io.h", ctypes.CDLL(io_path) if os.path.exists(io_path) else None)):
        if L is None:
            continue
        text = open(os.path.join(ROOT, "include", header)).read()
        names = set(re.findall(r"\b(fopis_[a-z_0-9]+)\s*\(", text))
        assert names
        for n in sorted(names):
            assert hasattr(L, n), f"{n} declared in {header} but not exported"


def _write_cfg(tmp_path, cfg, name="in.json"):
    # object paths resolve against the config directory
    for o in cfg.get("Object", {}).values():
        if not os.path.isabs(o["file_name"]):
            o["file_name"] = os.path.join(CONFIGS, o["file_name"])
    p = tmp_path / name
    p.write_text(json.dumps(cfg))
    return str(p)


def _host(path, **kw):
    from fopis_amd.host import Host
    return Host(path, no_device=True, **kw)


@pytest.mark.parametrize("name,nx", [("c1", 17), ("c2", 33), ("c3", 41)])
def test_desc_matches_oracle_normalisation(tmp_path, name, nx):
    cfg = load_config(name)
    h = _host(_write_cfg(tmp_path, cfg), nx_override=nx, ppc_override=4)
    g = refcpu.Grid(oracle_params(load_config(name), nx=nx, ppc=4))
    assert h.desc.nx == nx and h.desc.nspecies == g.S
    assert h.desc.eps0 == g.eps0()
    np.testing.assert_array_equal(np.array(h.desc.static_b[:]), g.static_b())
    for s in range(g.S):
        info = g.species_info(s)
        sd = h.desc.species[s]
        assert (sd.charge, sd.mass, sd.size) == (info["charge"], info["mass"], info["size"])
        np.testing.assert_array_equal(h.flux(s), g.calc_flux(s))


def test_initial_load_bitexact_with_objects(tmp_path):
    cfg = load_config("c2")
    h = _host(_write_cfg(tmp_path, cfg), nx_override=33, ppc_override=3)
    g = refcpu.Grid(oracle_params(load_config("c2"), nx=33, ppc=3))
    g.call("rc_load_initial_particles")
    for s in range(g.S):
        a = h.generate_initial(s)
        b = g.get_particles(s)
        assert len(a["x"]) == len(b["x"]) < 32 ** 3 * 3  # the cube's cells are empty
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(a[k], b[k])


def test_injection_bitexact(tmp_path):
    cfg = load_config("c2")
    h = _host(_write_cfg(tmp_path, cfg), nx_override=33, ppc_override=8)
    g = refcpu.Grid(oracle_params(load_config("c2"), nx=33, ppc=8, with_objects=False))
    counts = []
    for step in range(3):
        n0 = [g.num_particles(s) for s in range(g.S)]
        g.call("rc_inject")
        for s in range(g.S):
            a = h.generate_injection(s)
            b = g.get_particles(s)
            b = {k: v[n0[s]:] for k, v in b.items()}
            counts.append(len(a["x"]))
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                np.testing.assert_array_equal(a[k], b[k])
            # injected through the x-low face only, moving inward
            assert np.all(a["vx"] >= 0) and np.all(a["x"] <= a["vx"])
    assert min(counts) > 0


@pytest.mark.parametrize("obj,nx", [("golden/cube.obj", 9), ("golden/8x8x8_Sphere.obj", 17),
                                    ("cfg:cube16.obj", 33), ("cfg:lshape.obj", 65)])
def test_object_geometry_bitexact(tmp_path, obj, nx):
    path = os.path.join(CONFIGS, "objects", obj[4:]) if obj.startswith("cfg:") else os.path.join(ROOT, "tests", obj)
    cfg = load_config("c2")
    cfg["Object"]["Spacecraft"]["file_name"] = path
    h = _host(_write_cfg(tmp_path, cfg), nx_override=nx, ppc_override=1)
    g = refcpu.Grid(oracle_params(cfg, nx=nx, ppc=1))
    o = h.object(0)
    np.testing.assert_array_equal(o["mask"], g.cell_map(0))
    np.testing.assert_array_equal(o["cmat_ijk"], g.cmat_nodes(0))


def test_bad_inputs_raise(tmp_path):
    from fopis_amd.lib import FopisError
    cfg = load_config("c1")
    cfg["Environment"]["boundary"] = "DDDDDX"
    p = _write_cfg(tmp_path, cfg, "bad1.json")
    with pytest.raises(FopisError):
        _host(p)
    cfg = load_config("c1")
    cfg["Plasma"]["electron"]["temperature"] = 10.0  # v_th >= 0.4 violates checkPlasmaInfo
    with pytest.raises(FopisError, match="Plasma constraints"):
        _host(_write_cfg(tmp_path, cfg, "bad2.json"))
    cfg = load_config("c2")
    cfg["Object"]["Spacecraft"]["file_name"] = "objects/missing.obj"
    with pytest.raises(FopisError, match="does not exist"):
        _host(_write_cfg(tmp_path, cfg, "bad3.json"))


@pytest.mark.parametrize("shine", [(1.0, 0.0, 0.0), (0.0, 0.0, -1.0)])
def test_emission_bitexact(tmp_path, shine):
    """The host photoemission of GpuRootGrid (Spacecraft.cpp:829-900) against the
    oracle's emitParticlesFromObjects over three calls: the particles (moved by
    this step's position update, as the facade stages them), the surface charge
    change and the current change. Counts and currents are exact. Values are
    bit-identical except where the C library's sincos() and cos() round the
    same argument differently in the last bit (which of the two a compiler
    calls for sin/cos of one angle is its own choice), so they are held to
    4 ulp, with at most 0.5 % of them not bit-identical."""
    cfg = load_config("c2")
    c5 = load_config("c5")
    cfg["Plasma"]["photoelectron"] = c5["Plasma"]["photoelectron"]
    cfg["Object"]["Spacecraft"]["file_name"] = "objects/cube16.obj"
    cfg["Object"]["Spacecraft"]["emit_particles"] = c5["Object"]["Spacecraft"]["emit_particles"]
    cfg["Field"] = {"static_bfield": [0.0, 0.0, 0.0], "shine_vector": list(shine)}
    path = _write_cfg(tmp_path, cfg)
    h = _host(path, nx_override=33, load_particles=False, build_cmatrix=False, initialize=False)
    g = refcpu.Grid(oracle_params(cfg, nx=33))  # cfg's object path was made absolute above
    photo = 2
    for call in range(3):
        n0 = g.num_particles(photo)
        t0 = g.object_charge(1)[photo].copy()
        c0 = g.object_charge(2)[photo]
        g.call("rc_emit")
        a, dq, dcur = h.generate_emission(photo)
        b = g.get_particles(photo)
        b = {k: v[n0:] for k, v in b.items()}
        moved = {k: b[k] + b["v" + k] for k in ("x", "y", "z")}
        inside = np.ones(len(b["x"]), bool)
        for k in ("x", "y", "z"):
            inside &= (moved[k] >= 0.0) & (moved[k] <= 32.0)
        assert len(a["x"]) == int(inside.sum()) > 0
        got = np.concatenate([a[k] for k in ("x", "y", "z", "vx", "vy", "vz")])
        want = np.concatenate([moved[k][inside] for k in ("x", "y", "z")] +
                              [b["v" + k][inside] for k in ("x", "y", "z")])
        assert np.all(np.abs(got - want) <= 4 * np.spacing(np.abs(want)))
        assert np.mean(got != want) <= 0.005
        ref = g.object_charge(1)[photo] - t0
        np.testing.assert_allclose(dq, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        assert dcur == g.object_charge(2)[photo] - c0
