"""Shared test cases: reference-style inputs for the oracle and the matching
device description for libfopis_hip.so."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import refcpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIGS = os.path.join(ROOT, "fopis_amd", "configs")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_config(name: str) -> dict:
    with open(os.path.join(CONFIGS, name + ".json")) as f:
        return json.load(f)


def oracle_params(cfg: dict, nx: int | None = None, ppc: int | None = None, with_objects: bool = True,
                  bfield=None, post_loop: int = 1000) -> refcpu.Params:
    """refcpu.Params from a reference-schema config dict (species in reference id order)."""
    env = cfg["Environment"]
    n = nx or env["nx"]
    plasma = cfg["Plasma"]
    amb = sorted(k for k, v in plasma.items() if v["type"] == "ambient")
    photo = sorted(k for k, v in plasma.items() if v["type"] == "photoelectron")
    sps = []
    for name in amb + photo:
        v = plasma[name]
        sps.append(refcpu.species(v["type"], v["mass"], v["charge"], v["density"], v["temperature"],
                                  ppc or int(v["particle_per_cell"]), tuple(v.get("drift_velocity", (0, 0, 0))),
                                  tuple(int(b) for b in v.get("injection_axis", [0] * 6)),
                                  v.get("current_density", 0.0)))
    b = bfield if bfield is not None else cfg.get("Field", {}).get("static_bfield", [0.0, 0.0, 0.0])
    shine = cfg.get("Field", {}).get("shine_vector", [0.0, 0.0, 0.0])
    objs = []
    if with_objects:
        for name, o in sorted(cfg.get("Object", {}).items()):
            fn = o["file_name"]
            if not os.path.isabs(fn):
                fn = os.path.join(CONFIGS, fn)
            order = amb + photo
            emit = [order.index(k) for k in o.get("emit_particles", {}) if k in photo]
            objs.append({"file_name": fn, "is_potential_fixed": int(o.get("is_potential_fixed", False)),
                         "fixed_potential": o.get("fixed_potential", 0.0),
                         "initial_potential_offset": o.get("initial_potential_offset", 0.0),
                         "emit_species": emit[0] if emit else -1})
    return refcpu.params(n, n, n, env["dx"], env["dt"], sps, env["boundary"], tuple(b), tuple(shine), objs,
                         post_loop)


def device_desc(g: refcpu.Grid, precision=0, mg_precision=None, **kw):
    """Device description with the oracle's normalised species, eps0 and B."""
    from fopis_amd import lib as F
    sp = []
    for s in range(g.S):
        info = g.species_info(s)
        sp.append((info["charge"], info["mass"], info["size"]))
    return F.make_desc(g.nx, g.ny, g.nz, sp, g.eps0(), tuple(g.static_b()), "DDDDDD", precision,
                       mg_precision, **kw)


def canon(d: dict, drop_nan: bool = True) -> np.ndarray:
    """Particles as an (n, 6) array sorted lexicographically by x, y, z, vx, vy, vz."""
    a = np.stack([np.asarray(d[k], np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")], axis=1)
    if "valid" in d:
        a = a[np.asarray(d["valid"]) != 0]
    if drop_nan:
        a = a[~np.isnan(a[:, 0])]
    order = np.lexsort(a.T[::-1])
    return a[order]


def dst_poisson(b: np.ndarray) -> np.ndarray:
    """Exact solution of (6 phi - sum of 6 neighbours) = b with zero Dirichlet faces (DST-I)."""
    from scipy.fft import dstn, idstn
    n0, n1, n2 = b.shape
    k0 = np.arange(1, n0 + 1)[:, None, None]
    k1 = np.arange(1, n1 + 1)[None, :, None]
    k2 = np.arange(1, n2 + 1)[None, None, :]
    lam = (6.0 - 2.0 * np.cos(np.pi * k0 / (n0 + 1)) - 2.0 * np.cos(np.pi * k1 / (n1 + 1))
           - 2.0 * np.cos(np.pi * k2 / (n2 + 1)))
    return idstn(dstn(b, type=1) / lam, type=1)
