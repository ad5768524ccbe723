"""FP32 step against the FP64 step on the same seeded input.

Runs a configuration twice from the same device-drawn plasma: once with FP64
particles and V-cycles, once with FP32 particles, FP32 deposit arithmetic and
FP32 V-cycles (the FP64 defect correction, conductor algebra and reductions
stay FP64 in both). Reports the error of the FP32 run against the FP64 run for
rho, phi, the particle phase-space moments and the energies, right after
initialisation and after --steps steps, as one JSON line.

  python -m fopis_amd.precision_check --config fopis_amd/configs/c3.json --steps 10
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from fopis_amd import lib as F  # noqa: E402
from fopis_amd.host import Host  # noqa: E402


def log(msg):
    print(f"[precision_check] {msg}", file=sys.stderr, flush=True)


def with_precision(path: str, prec: str) -> str:
    with open(path) as f:
        cfg = json.load(f)
    cfg.setdefault("Build", {})
    cfg["Build"]["precision"] = prec
    cfg["Build"]["mg_precision"] = prec
    base = os.path.dirname(os.path.abspath(path))
    for o in cfg.get("Object", {}).values():
        o["file_name"] = os.path.join(base, o["file_name"])
    fd, out = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(cfg, f)
    return out


def snapshot(h: Host) -> dict:
    ctx = h.ctx
    ctx.synchronize()
    snap = {"rho": ctx.get_field(F.RHO, 0), "phi": ctx.get_field(F.PHI),
            "moments": [ctx.moments(s) for s in range(h.S)], "energy": ctx.energy()}
    return snap


def run(path: str, prec: str, steps: int, seed: int, ppc: int, cache: str):
    cfg = with_precision(path, prec)
    t0 = time.perf_counter()
    h = Host(cfg, device_loader=True, device_seed=seed, ppc_override=ppc, build_cmatrix=True, initialize=True,
             cmatrix_cache=cache)
    os.unlink(cfg)
    log(f"{prec}: setup {time.perf_counter() - t0:.1f} s")
    s0 = snapshot(h)
    for i in range(steps):
        h.main_loop_es()
    sN = snapshot(h)
    log(f"{prec}: {steps} steps done")
    h.close()
    return s0, sN


def field_err(a32, a64):
    d = a32 - a64
    return {"rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(a64), 1e-300)),
            "max_abs_over_max": float(np.max(np.abs(d)) / max(np.max(np.abs(a64)), 1e-300))}


def compare(s32, s64):
    out = {"rho": field_err(s32["rho"], s64["rho"]), "phi": field_err(s32["phi"], s64["phi"]), "species": []}
    for m32, m64 in zip(s32["moments"], s64["moments"]):
        sig = np.sqrt(np.maximum(m64["var_v"], 1e-300))
        out["species"].append({
            "count_64": int(m64["count"]), "count_diff": int(m32["count"] - m64["count"]),
            "mean_v_diff_over_sigma": [float(x) for x in (m32["mean_v"] - m64["mean_v"]) / sig],
            "var_v_rel_diff": [float(x) for x in (m32["var_v"] - m64["var_v"]) / np.maximum(m64["var_v"], 1e-300)],
            "mean_x_diff": [float(x) for x in (m32["mean_x"] - m64["mean_x"])],
        })
    e32, e64 = s32["energy"], s64["energy"]
    out["particle_energy_rel_diff"] = float((e32[0] - e64[0]) / e64[0])
    out["field_energy_rel_diff"] = float((e32[1] - e64[1]) / e64[1])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default=os.path.join(HERE, "configs", "c3.json"))
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--seed", type=int, default=20251202)
    ap.add_argument("--ppc", type=int, default=0, help="particles per cell override (0 = config)")
    ap.add_argument("--cmatrix-cache", default="", help="capacity-matrix cache directory shared by both runs")
    args = ap.parse_args(argv)
    cache = args.cmatrix_cache or tempfile.mkdtemp(prefix="fopis_cmat_")
    a0, aN = run(args.config, "fp64", args.steps, args.seed, args.ppc, cache)
    b0, bN = run(args.config, "fp32", args.steps, args.seed, args.ppc, cache)
    if not args.cmatrix_cache:
        shutil.rmtree(cache, ignore_errors=True)
    res = {"config": os.path.basename(args.config), "steps": args.steps, "ppc": args.ppc or None,
           "after_initialize": compare(b0, a0), f"after_{args.steps}_steps": compare(bN, aN)}
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
