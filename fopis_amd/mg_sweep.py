"""Multigrid parameter sweep on a configuration's own step data.

Runs the configuration (objects removed, particles drawn on the device) for
a few steps, keeps phi of step n and rho of step n+1, and for each
(omega, pre, post) solves rho_{n+1} to the configured tolerance from phi_n
(the step's warm start) and from zero, reporting V-cycles and milliseconds.
One JSON line per setting.

  python -m fopis_amd.mg_sweep --config fopis_amd/configs/c3.json \
      --settings 1.5:2:2 1.3:2:2 1.25:2:2 1.2:2:2
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from fopis_amd import lib as F  # noqa: E402
from fopis_amd.host import Host  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default=os.path.join(HERE, "configs", "c2.json"))
    ap.add_argument("--settings", nargs="+", default=["1.5:2:2", "1.3:2:2", "1.25:2:2", "1.2:2:2"],
                    help="omega:pre:post triples")
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args(argv)
    with open(args.config) as f:
        cfg = json.load(f)
    cfg["Object"] = {}
    fd, tmp = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(cfg, f)
    h = Host(tmp, device_loader=True, device_seed=args.seed, build_cmatrix=False, initialize=True,
             device_injection=True)
    os.unlink(tmp)
    for _ in range(2):
        h.main_loop_es()
    phi_n = h.ctx.get_field(F.PHI)
    h.main_loop_es()
    rho_n1 = h.ctx.get_field(F.RHO, 0)
    base = copy.copy(h.desc)
    h.close()
    out = []
    for s in args.settings:
        om, pre, post = s.split(":")
        d = copy.copy(base)
        d.mg_omega, d.mg_pre, d.mg_post = float(om), int(pre), int(post)
        ctx = F.Context(d)
        res = {"config": os.path.basename(args.config), "omega": float(om), "pre": int(pre), "post": int(post)}
        for start in ("warm", "zero"):
            ctx.set_field(F.RHO, rho_n1)
            ctx.set_field(F.PHI, phi_n if start == "warm" else np.zeros_like(phi_n))
            ctx.synchronize()
            ctx.reset_timers()
            cyc, rr = ctx.solve_poisson()
            res[start] = {"cycles": cyc, "ms": ctx.phase_time("poisson")[0], "relres": rr}
        ctx.close()
        print(json.dumps(res), flush=True)
        out.append(res)
    return out


if __name__ == "__main__":
    main()
