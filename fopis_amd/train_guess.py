"""Train the learned Poisson initial guess on pairs the solver generates.

Recipe (the 128^3 learned-guess workload): rho is the deposit of a plasma
of the given configuration (its species, densities and temperatures, drawn
on the device; one draw serves --plasma-every pairs) plus 1-3 random
axis-aligned conductor boxes of 8-32 cells per side whose surface nodes carry a seeded charge;
phi is the multigrid solution of that rho to relative residual 1e-8 from
zero. Pairs [0, ntrain) train the network, pairs after that are held out.
Training, inference and the pair store all run in libfopis_hip.so.

Writes <out>.npz (params float32, out_scale, grid, and the recipe) and prints
one JSON line with the losses and, on held-out problems, multigrid cycles and
milliseconds from zero and from the network's guess. --pairs-out writes the
generated (rho_hat, mask, phi_hat) pairs to an HDF5 file and --pairs-in trains
on such a file instead of generating (Guess.export_pairs / import_pairs).

  python -m fopis_amd.train_guess --config fopis_amd/configs/c2.json \
      --ntrain 2000 --nheld 200 --epochs 8 --out fopis_amd/weights/guess_c4
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from fopis_amd import lib as F  # noqa: E402
from fopis_amd.guess import Guess  # noqa: E402
from fopis_amd.host import Host  # noqa: E402


def log(msg):
    print(f"[train_guess] {msg}", file=sys.stderr, flush=True)


def plasma_config(path: str) -> str:
    """The configuration without its objects (the boxes stand in for them)."""
    with open(path) as f:
        cfg = json.load(f)
    cfg["Object"] = {}
    fd, out = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(cfg, f)
    return out


def draw_problem(rng, nx: int, rho_rms: float, smin: int, smax: int):
    """1-3 boxes (glue cell ranges) and their surface charge per node."""
    boxes, charges = [], []
    for _ in range(int(rng.integers(1, 4))):
        s = rng.integers(smin, smax + 1, 3)
        lo = [int(rng.integers(2, nx - 1 - int(s[a]))) for a in range(3)]
        boxes.append(lo + [lo[a] + int(s[a]) for a in range(3)])
        charges.append(float(rng.choice([-1.0, 1.0]) * rho_rms * 10.0 ** rng.uniform(-0.5, 1.0)))
    return boxes, charges


class PairMaker:
    """Problem i = plasma draw i // every (a fresh device load of the
    configuration's species, deposited) + its own seeded boxes."""

    def __init__(self, host: Host, guess: Guess, seed: int, smin: int, smax: int, every: int):
        self.h, self.g, self.ctx = host, guess, host.ctx
        self.seed, self.smin, self.smax, self.every = seed, smin, smax, max(1, every)
        self.zero = np.zeros(self.ctx.field_shape(F.PHI))
        self.plasma_id, self.plasma = None, None
        self._plasma(0)
        n = self.ctx.nx
        self.rho_rms = float(np.sqrt(np.mean(self.plasma[2:n, 2:n, 2:n] ** 2)))

    def _plasma(self, pid: int):
        if pid != self.plasma_id:
            self.h.reload(self.seed + 7 * pid + 1)
            self.ctx.update_rho()
            self.plasma = self.ctx.get_field(F.RHO)
            self.plasma_id = pid

    def setup(self, i: int):
        """rho and the boxes of problem i (deterministic in i); phi = 0."""
        rng = np.random.default_rng(self.seed * 1_000_003 + i)
        self._plasma(i // self.every)
        self.ctx.set_field(F.RHO, self.plasma)
        boxes, charges = draw_problem(rng, self.ctx.nx, self.rho_rms, self.smin, self.smax)
        self.g.synth_boxes(boxes, charges, seed=self.seed + i)
        self.ctx.set_field(F.PHI, self.zero)

    def solve(self):
        self.ctx.reset_timers()
        cyc, rr = self.ctx.solve_poisson()
        ms = self.ctx.phase_time("poisson")[0]
        return cyc, rr, ms


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default=os.path.join(HERE, "configs", "c2.json"))
    ap.add_argument("--ntrain", type=int, default=2000)
    ap.add_argument("--nheld", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-end", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=20251202)
    ap.add_argument("--box-min", type=int, default=8)
    ap.add_argument("--box-max", type=int, default=32)
    ap.add_argument("--eval-problems", type=int, default=16, help="held-out problems for the cycle comparison")
    ap.add_argument("--plasma-every", type=int, default=10, help="pairs that share one plasma draw")
    ap.add_argument("--loss", type=int, default=1,
                    help="0: potential MSE; 1: the multigrid's starting residual (sets its cycle count)")
    ap.add_argument("--init", default="", help="start from these weights (.npz)")
    ap.add_argument("--out", default="")
    ap.add_argument("--pairs-out", default="", help="also write the generated pairs to this HDF5 file")
    ap.add_argument("--pairs-in", default="",
                    help="train on the pairs of this HDF5 file (written by --pairs-out) instead of generating them")
    ap.add_argument("--correction", action="store_true",
                    help="train the correction of the warm start (guess mode 2) on pairs recorded from PIC runs")
    ap.add_argument("--steps-per-run", type=int, default=250, help="--correction: pairs recorded per run")
    ap.add_argument("--warm-steps", type=int, default=3, help="--correction: steps before recording")
    ap.add_argument("--eval-runs", type=int, default=2, help="--correction: held-out runs")
    ap.add_argument("--eval-steps", type=int, default=5, help="--correction: steps per mode and run")
    ap.add_argument("--precision", default="fp32", help="--correction: particle precision of the runs")
    args = ap.parse_args(argv)
    if args.correction:
        return main_correction(args)

    cfg = plasma_config(args.config)
    t0 = time.perf_counter()
    h = Host(cfg, device_loader=True, device_seed=args.seed, build_cmatrix=False, initialize=False)
    os.unlink(cfg)
    ctx = h.ctx
    g = Guess(ctx, seed=args.seed)
    g.set_loss(args.loss)
    if args.init:
        z = np.load(args.init, allow_pickle=False)
        g.set_params(z["params"], float(z["out_scale"]))
    pm = PairMaker(h, g, args.seed, args.box_min, args.box_max, args.plasma_every)
    log(f"grid {g.grid}, {g.nparams} parameters, rho rms {pm.rho_rms:.4g}, setup {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    cycles = []
    if args.pairs_in:
        from fopis_amd.h5 import H5File

        with H5File(args.pairs_in, "r") as f:
            nfile = int(f.read("count", np.int32))
        if nfile < args.ntrain + args.nheld:
            raise ValueError(f"{args.pairs_in} holds {nfile} pairs, {args.ntrain} + {args.nheld} asked for")
        ntot = nfile
        g.dataset_alloc(ntot)
        g.import_pairs(args.pairs_in)
        log(f"read {ntot} pairs from {args.pairs_in}")
    else:
        ntot = args.ntrain + args.nheld
        g.dataset_alloc(ntot)
        for i in range(ntot):
            pm.setup(i)
            cyc, rr, _ = pm.solve()
            if not rr <= 1e-8:
                raise RuntimeError(f"pair {i}: solve stopped at relres {rr}")
            cycles.append(cyc)
            g.dataset_store(i)
            if (i + 1) % max(1, ntot // 10) == 0:
                log(f"pairs {i + 1}/{ntot} ({time.perf_counter() - t0:.1f} s)")
        if args.pairs_out:
            g.export_pairs(args.pairs_out, 0, ntot)
            log(f"wrote {ntot} pairs to {args.pairs_out}")
    gen_s = time.perf_counter() - t0

    held = (args.ntrain, args.ntrain + args.nheld) if args.nheld else (0, args.ntrain)
    held_before = g.eval(*held)
    t0 = time.perf_counter()
    losses = []
    for e in range(args.epochs):
        # one cosine over all epochs, taken one epoch per call
        f0, f1 = e / args.epochs, (e + 1) / args.epochs
        lr0 = args.lr_end + (args.lr - args.lr_end) * 0.5 * (1 + math.cos(math.pi * f0))
        lr1 = args.lr_end + (args.lr - args.lr_end) * 0.5 * (1 + math.cos(math.pi * f1))
        losses.append(float(g.train(args.ntrain, 1, lr0, lr1, seed=args.seed + e)[0]))
        log(f"epoch {e + 1}/{args.epochs}: train loss {losses[-1]:.5g} ({time.perf_counter() - t0:.1f} s)")
    train_s = time.perf_counter() - t0
    held_after = g.eval(*held)
    params, k = g.get_params()

    # held-out problems: multigrid from zero vs from the network
    rz, rg = [], []
    for j in range(min(args.eval_problems, max(1, args.nheld))):
        i = args.ntrain + j
        pm.setup(i)
        rz.append(pm.solve())
        pm.setup(i)
        ctx.reset_timers()
        g.apply()
        gms = ctx.phase_time("guess")[0]
        cyc, rr, ms = pm.solve()
        rg.append((cyc, rr, ms, gms))
    res = {
        "grid": list(g.grid), "params": int(g.nparams), "out_scale": k,
        "pairs": {"train": args.ntrain, "held_out": args.nheld, "generate_s": gen_s,
                  "cycles_from_zero_mean": float(np.mean(cycles)) if cycles else None},
        "loss": {"train_first_epoch": losses[0] if losses else None, "train_last_epoch": losses[-1] if losses else None,
                 "held_out_before": held_before, "held_out_after": held_after},
        "train_s": train_s, "steps_per_s": args.epochs * args.ntrain / max(train_s, 1e-9),
        "eval": {
            "problems": len(rz),
            "cycles_from_zero": float(np.mean([r[0] for r in rz])),
            "ms_from_zero": float(np.mean([r[2] for r in rz])),
            "cycles_from_guess": float(np.mean([r[0] for r in rg])),
            "ms_from_guess": float(np.mean([r[2] for r in rg])),
            "guess_ms": float(np.mean([r[3] for r in rg])),
        },
    }
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        np.savez(args.out, params=params, out_scale=np.float64(k), grid=np.array(g.grid, np.int32),
                 recipe=np.array(json.dumps({"config": os.path.basename(args.config), "ntrain": args.ntrain,
                                             "epochs": args.epochs, "seed": args.seed, "loss": args.loss,
                                             "boxes": [args.box_min, args.box_max]})))
        res["weights"] = args.out + ".npz"
    h.close()
    print(json.dumps(res), flush=True)
    return res


def box_config(base: str, boxes, tmpdir: str, precision: str) -> str:
    """The configuration with its object replaced by a perfect conductor made
    of the union of `boxes` (cell ranges [lo, hi) per axis), written as a
    voxel .obj in the reference's coordinates (vertices relative to nx/2,
    ObjectUtils.cpp:57-59)."""
    from fopis_amd.voxel_obj import voxels_from_boxes, write_obj

    with open(base) as f:
        cfg = json.load(f)
    nx = cfg["Environment"]["nx"]
    lo = [min(b[0][q] for b in boxes) for q in range(3)]
    hi = [max(b[1][q] for b in boxes) for q in range(3)]
    obj = os.path.join(tmpdir, "boxes.obj")
    write_obj(obj, voxels_from_boxes(boxes, lo, hi), [lo[q] - nx // 2 for q in range(3)])
    spec = {"file_name": obj, "surface_type": "perfect", "emit_particles": {}, "materials": {},
            "plot_potential_mapping_width": 0, "is_potential_fixed": False, "fixed_potential": 0.0,
            "initial_potential_offset": 0.0}
    cfg["Object"] = {"Spacecraft": spec}
    cfg.setdefault("Build", {})["precision"] = precision
    out = os.path.join(tmpdir, "run.json")
    with open(out, "w") as f:
        json.dump(cfg, f)
    return out


def draw_boxes(rng, ncell: int, smin: int, smax: int):
    """1-3 axis-aligned boxes of smin-smax cells per side, 4 cells or more from the faces."""
    out = []
    for _ in range(int(rng.integers(1, 4))):
        s = rng.integers(smin, smax + 1, 3)
        lo = [int(rng.integers(4, ncell - 4 - int(s[q]))) for q in range(3)]
        out.append((lo, [lo[q] + int(s[q]) for q in range(3)]))
    return out


def main_correction(args):
    """The warm-start recipe: pairs are (the residual charge of the step's
    warm start, the potential the first solve adds to it) recorded from PIC
    runs of the configuration's plasma around 1-3 random box conductors
    (guess mode 3); the network learns the correction (mode 2). Held-out runs
    compare the first solve of each step from the warm start against the
    same solve from the warm start plus the network's correction."""
    import shutil

    t0 = time.perf_counter()
    tmp = tempfile.mkdtemp(prefix="fopis_corr_")
    rng = np.random.default_rng(args.seed)
    ntot = args.ntrain + args.nheld
    train = None
    try:
        slot = 0
        run = 0
        while slot < ntot:
            with open(args.config) as f:
                ncell = json.load(f)["Environment"]["nx"] - 1
            boxes = draw_boxes(rng, ncell, args.box_min, args.box_max)
            cfg = box_config(args.config, boxes, tmp, args.precision)
            log(f"run {run}: boxes {boxes}")
            h = Host(cfg, device_loader=True, device_seed=args.seed + 101 * run, build_cmatrix=True, initialize=True,
                     device_injection=True)
            gr = Guess(h.ctx, seed=args.seed)
            if train is None:
                # the network and the dataset live in the first run's grid size, in a context of their own
                tctx = F.Context(F.make_desc(h.ctx.nx, h.ctx.ny, h.ctx.nz, [(-1.0, 1.0, 1.0)], eps0=1.0,
                                             mg_precision=F.FP32))
                train = Guess(tctx, seed=args.seed)
                train.set_loss(args.loss)
                if args.init:
                    z = np.load(args.init, allow_pickle=False)
                    train.set_params(z["params"], float(z["out_scale"]))
                train.dataset_alloc(ntot)
            take = min(args.steps_per_run, ntot - slot)
            gr.dataset_alloc(take)
            for _ in range(args.warm_steps):
                h.main_loop_es()
            gr.set_mode(3)
            while gr.recorded() < take:
                h.main_loop_es()
            gr.set_mode(0)
            for i in range(take):
                train.dataset_set(slot + i, *gr.dataset_get(i))
            slot += take
            log(f"run {run}: {len(boxes)} boxes, {take} pairs ({slot}/{ntot}, {time.perf_counter() - t0:.1f} s)")
            h.close()
            run += 1
        gen_s = time.perf_counter() - t0
        held = (args.ntrain, ntot) if args.nheld else (0, args.ntrain)
        held_before = train.eval(*held)
        t0 = time.perf_counter()
        losses = []
        best = (held_before, train.get_params())
        for e in range(args.epochs):
            f0, f1 = e / args.epochs, (e + 1) / args.epochs
            lr0 = args.lr_end + (args.lr - args.lr_end) * 0.5 * (1 + math.cos(math.pi * f0))
            lr1 = args.lr_end + (args.lr - args.lr_end) * 0.5 * (1 + math.cos(math.pi * f1))
            losses.append(float(train.train(args.ntrain, 1, lr0, lr1, seed=args.seed + e)[0]))
            ho = train.eval(*held)
            if ho < best[0]:
                best = (ho, train.get_params())
            log(f"epoch {e + 1}/{args.epochs}: train loss {losses[-1]:.5g}, held-out {ho:.5g} "
                f"({time.perf_counter() - t0:.1f} s)")
        train_s = time.perf_counter() - t0
        held_after, (params, k) = best
        train.ctx.close()

        # held-out runs: the first solve of each step from the warm start and from the warm start + correction
        ev = {"plain": [], "learned": []}
        for j in range(args.eval_runs):
            boxes = draw_boxes(rng, ncell, args.box_min, args.box_max)
            cfg = box_config(args.config, boxes, tmp, args.precision)
            h = Host(cfg, device_loader=True, device_seed=args.seed + 7919 * (j + 1), build_cmatrix=True,
                     initialize=True, device_injection=True)
            g = Guess(h.ctx, seed=args.seed)
            g.set_params(params, k)
            for _ in range(args.warm_steps):
                h.main_loop_es()
            for mode, key in ((0, "plain"), (2, "learned"), (0, "plain"), (2, "learned")):
                g.set_mode(mode)
                h.main_loop_es()
                h.ctx.synchronize()
                h.ctx.reset_timers()
                for _ in range(args.eval_steps):
                    h.main_loop_es()
                h.ctx.synchronize()
                st = F.solver_stats(h.ctx)
                ev[key].append({"first_solve_ms": h.ctx.phase_time("poisson_first")[0],
                                "guess_ms": h.ctx.phase_time("guess")[0] if mode else 0.0,
                                "cycles": st["first_cycles"] / args.eval_steps,
                                "start_relres": st["start_relres_first"]})
            h.close()
        summary = {key: {q: float(np.mean([r[q] for r in v])) for q in v[0]} for key, v in ev.items()}
        res = {"recipe": "correction", "grid": list(train.grid), "params": int(train.nparams), "out_scale": k,
               "pairs": {"train": args.ntrain, "held_out": args.nheld, "runs": run, "generate_s": gen_s},
               "loss": {"train_first_epoch": losses[0] if losses else None,
                        "train_last_epoch": losses[-1] if losses else None,
                        "held_out_before": held_before, "held_out_after": held_after},
               "train_s": train_s, "eval": summary,
               "speedup": summary["plain"]["first_solve_ms"] / max(summary["learned"]["first_solve_ms"], 1e-9)}
        if args.out:
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            np.savez(args.out, params=params, out_scale=np.float64(k), grid=np.array(train.grid, np.int32),
                     recipe=np.array(json.dumps({"config": os.path.basename(args.config), "recipe": "correction",
                                                 "ntrain": args.ntrain, "epochs": args.epochs, "seed": args.seed,
                                                 "loss": args.loss, "boxes": [args.box_min, args.box_max],
                                                 "init": os.path.basename(args.init)})))
            res["weights"] = args.out + ".npz"
        print(json.dumps(res), flush=True)
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
