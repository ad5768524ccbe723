"""Benchmark of the MI355X ES PIC step (one GPU).

Default workload (C3): 256^3 cells, 64 particles per cell for each of
electrons and ions (2 x 1.07e9 superparticles, FP32 positions and
velocities), a floating 32^3-cell cube conductor handled by the capacitance
matrix, B0 = 0.2 along z (full Boris rotation), upstream x-face injection,
multigrid Poisson (FP32 V(2,2) cycles, FP64 defect correction) to relative
residual 1e-8 (two solves per step because of the conductor). Particles are
drawn on the device (Philox Maxwellian, synthetic); boundary injection keeps
the reference's per-face counts and residues on the host and draws the
particles on the device (--injection host draws them on the host from the
reference's mt19937 streams instead). --workload c2 runs the 128^3 FP64
case, c5 the 256^3 L-shaped spacecraft with photoelectrons.

A step is one RootGrid::mainLoopES. `value` is particles/s of the push+deposit
part of the step (gather, Boris, move, absorb, count, CIC deposit, re-binning
and the window reduction into rho), timed with HIP events on the library's
stream inside the timed steps; ms_per_step is the whole step by wall clock.
The roofline object is k_tile's achieved algorithmic rate against the
nominal 8 TB/s; hbm_copy_gbs is the device-to-device copy rate measured in
the same run (the practical streaming ceiling) and roofline.frac_of_copy the
kernel's rate against it. cpu_baseline times the oracle's push + move +
deposit on every core the process may use; poisson_cpu_baseline times the
reference's SOR on the right-hand side of the GPU's first solve of a step,
warm-started from the previous step's phi.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s (push+CIC deposit); Poisson ms to rel-res 1e-8, learned vs plain"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c2": {"json": "c2.json", "bytes_per_particle": {0: 96, 1: 48},
           "name": "C2: 128^3 cells, 64 ppc x 2 species, floating 16^3 cube, x-face injection, FP64"},
    "c3": {"json": "c3.json", "bytes_per_particle": {0: 96, 1: 48},
           "name": "C3: 256^3 cells, 64 ppc x 2 species, floating 32^3 cube, B0=0.2 z, FP32 push/deposit"},
    "c5": {"json": "c5.json", "bytes_per_particle": {0: 96, 1: 48},
           "name": "C5: 256^3 cells, 64 ppc x 2 species + photoelectrons (8 per lit surface node per step, vth 0.3) "
                   "from the -x face of an L-shaped floating spacecraft (48x16x16 + 16x48x16), B0=0.2 z, FP32"},
    "c1": {"json": "c1.json", "bytes_per_particle": {0: 96, 1: 48},
           "name": "C1: 32^3 cells, 20 ppc x 2 species, no object, FP64"},
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info() -> dict:
    """The host the CPU baseline runs on: logical CPUs, those this process
    may run on, the cgroup CPU quota, the CPU model and the RAM."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
        info["cgroup_cpus"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    try:
        with open("/proc/cpuinfo") as f:
            info["model"] = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        info["model"] = None
    try:
        with open("/proc/meminfo") as f:
            kb = next(int(l.split()[1]) for l in f if l.startswith("MemTotal"))
        info["ram_gb"] = round(kb / 2 ** 20, 1)
    except (OSError, StopIteration):
        info["ram_gb"] = None
    env = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads_env"] = int(env) if env and env.isdigit() else None
    return info


def cpu_threads(info: dict) -> int:
    """Every core this process may use: the affinity set, bounded by the
    cgroup quota and by OMP_NUM_THREADS when the box sets it (its CPU share)."""
    n = info["affinity"] or 1
    if info["cgroup_cpus"]:
        n = min(n, max(1, int(info["cgroup_cpus"])))
    if info["omp_num_threads_env"]:
        n = min(n, info["omp_num_threads_env"])
    return n


def cpu_baseline(cfg_name: str, threads: int, info: dict, ppc: int = 2, samples: int = 5) -> dict:
    """The oracle (CPU restatement of the reference step: AoS FP64 particles,
    OpenMP over particles for the push and move, over species for the
    deposit) timed on push + move + deposit at the workload's grid with a
    reduced particle count; median of `samples` steps."""
    from oracle import refcpu
    from tests.cases import load_config, oracle_params

    refcpu.set_threads(threads)
    cfg = load_config(cfg_name)
    g = refcpu.Grid(oracle_params(cfg, ppc=ppc))
    t0 = time.perf_counter()
    g.call("rc_load_initial_particles")
    load_s = time.perf_counter() - t0
    n = sum(g.num_particles(s) for s in range(g.S))
    ts = []
    for _ in range(samples):
        t0 = time.perf_counter()
        g.call("rc_update_velocity")
        g.call("rc_update_position")
        g.call("rc_deposit_only")
        ts.append(time.perf_counter() - t0)
    med = sorted(ts)[len(ts) // 2]
    nx = cfg["Environment"]["nx"]
    del g
    return {"value": n / med, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"oracle push+move+deposit (reference algorithm and layout: FP64 AoS, OpenMP over particles, "
                      f"deposit parallel over species) on the {nx - 1}^3 grid with the conductor, {ppc} ppc per "
                      f"species ({n} particles), median of {samples} steps",
            "step_s": ts, "load_s": load_s, "host": info}


def poisson_cpu_baseline(desc, cfg_name: str, phi_prev, rho_new, threads: int, gpu_ms: float,
                         gpu_cycles: float) -> dict:
    """The reference's Poisson solver (zebra SOR, omega = 2/(1 + sin(pi/n)),
    RootGridField.cpp:65-264) in the oracle on the same right-hand side the
    GPU's first solve of a step had, warm-started from the previous step's
    phi, as the reference warm-starts: (i) with the reference's own stop (the
    pointwise relative residual below 1e-7 checked every 10 sweeps, at most
    1000 sweeps), (ii) to this build's ||b - A phi||_2 / ||b||_2 <= 1e-8."""
    from oracle import refcpu
    from tests.cases import load_config, oracle_params

    refcpu.set_threads(threads)
    g = refcpu.Grid(oracle_params(load_config(cfg_name), ppc=1, with_objects=False))
    out = {"unit": "ms", "cores": threads, "kind": "port", "gpu_first_solve_ms": gpu_ms,
           "gpu_first_solve_cycles": gpu_cycles}
    g.set_field(refcpu.RHO, rho_new, 0)
    for key, run in (("reference_stop", lambda: g.call("rc_solve_poisson")),
                     ("relres_1e-8", lambda: g.call("rc_solve_poisson_relres", 1e-8, 100000))):
        g.set_field(refcpu.PHI, phi_prev)
        t0 = time.perf_counter()
        sweeps = run()
        dt = time.perf_counter() - t0
        out[key] = {"ms": dt * 1e3, "sweeps": int(sweeps), "relres": float(g.L.rc_relres(g.h))}
    out["value"] = out["relres_1e-8"]["ms"]
    out["gpu_speedup_to_1e-8"] = out["relres_1e-8"]["ms"] / gpu_ms if gpu_ms > 0 else None
    del g
    return out


def learned_leg(steps: int, weights: str, cmatrix_cache: str = "") -> dict:
    """The learned-guess workload (C4: 128^3 cells, the C2 plasma around its
    floating 16^3 cube, FP32 particles): the first Poisson solve of each step
    from the warm start against the same solve from the warm start plus the
    network's correction (guess mode 2; guess mode 1, the network's potential
    alone, for weights trained on rho -> phi), both over `steps` steps, with
    guess inference included and the relative residual each start leaves."""
    import numpy as np

    from fopis_amd import lib as F
    from fopis_amd.guess import Guess
    from fopis_amd.host import Host

    z = np.load(weights, allow_pickle=False)
    recipe = json.loads(str(z["recipe"])) if "recipe" in z.files else {}
    learned_mode = 2 if recipe.get("recipe") == "correction" else 1
    cfg_path = os.path.join(ROOT, "fopis_amd", "configs", "c2.json")
    with open(cfg_path) as f:
        cfg = json.load(f)
    cfg.setdefault("Build", {})["precision"] = "fp32"
    for o in cfg.get("Object", {}).values():
        o["file_name"] = os.path.join(os.path.dirname(cfg_path), o["file_name"])
    import tempfile

    fd, tmp = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(cfg, f)
    h = Host(tmp, device_loader=True, device_seed=20251203, build_cmatrix=True, initialize=True,
             cmatrix_cache=cmatrix_cache, device_injection=True)
    os.unlink(tmp)
    ctx = h.ctx
    out = {"workload": "C4: 128^3 cells, the C2 plasma (64 ppc x 2 species, FP32) around its floating 16^3 cube",
           "weights": os.path.relpath(weights, ROOT), "mode": learned_mode,
           "trained_grid": [int(v) for v in z["grid"]]}
    g = Guess(ctx, seed=1)
    g.set_params(z["params"], float(z["out_scale"]))
    for _ in range(3):
        h.main_loop_es()
    for mode, key in ((0, "plain"), (learned_mode, "learned")):
        g.set_mode(mode)
        h.main_loop_es()
        ctx.synchronize()
        ctx.reset_timers()
        for _ in range(steps):
            h.main_loop_es()
        ctx.synchronize()
        st = F.solver_stats(ctx)
        out[key] = {"first_solve_ms": ctx.phase_time("poisson_first")[0],
                    "first_solve_cycles": st["first_cycles"] / steps,
                    "start_relres": st["start_relres_first"],
                    "guess_ms": ctx.phase_time("guess")[0] if mode else 0.0}
    g.set_mode(0)
    h.close()
    out["speedup"] = out["plain"]["first_solve_ms"] / max(out["learned"]["first_solve_ms"], 1e-9)
    out["cycles_saved"] = out["plain"]["first_solve_cycles"] - out["learned"]["first_solve_cycles"]
    return out


def copy_ceiling_gbs(reps: int = 5) -> float:
    """Device-to-device copy bandwidth (read + write bytes per second) of six
    512 MiB FP32 arrays, the shape of the particle streams: the HBM rate a
    plain streaming kernel reaches on this device, for reading the roofline
    fraction against the nominal 8 TB/s."""
    import torch

    n = 1 << 27
    src = [torch.ones(n, dtype=torch.float32, device="cuda") for _ in range(6)]
    dst = [torch.empty_like(a) for a in src]
    for a, b in zip(src, dst):
        b.copy_(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        for a, b in zip(src, dst):
            b.copy_(a)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * 6 * n * 4 / (ms * 1e-3) / 1e9


def pmc_traffic(workload: str):
    p = os.path.join(ROOT, "profiles", f"tile_kernel_{workload}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def poisson_roofline(desc, ms, cycles, updates=None, pair_passes=0.0):
    """Algorithmic HBM bytes of one Poisson solve over its measured time. Per
    fine-grid unknown: a V(2,2) cycle moves 83 B with FP32 levels (166 B with
    FP64: four smoothing sweeps of 12 B plus residual, restriction and
    prolongation, and 8/7 over the coarser levels), each FP64 update of the
    defect correction 40 B (one per cycle, or one per pair of cycles sharing a
    right-hand side), each residual pass between the cycles of a pair 8 B
    (16 B with FP64 levels), and the two residuals formed from scratch (start
    and end) 28 B each. updates and pair_passes are per solve."""
    n = (desc.nx - 2) * (desc.ny - 2) * (desc.nz - 2)
    vb = 83 if desc.mg_precision == 1 else 166
    pb = 8 if desc.mg_precision == 1 else 16
    if updates is None:
        updates = cycles
    nbytes = n * (cycles * vb + updates * 40 + pair_passes * pb + 2 * 28)
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "algorithmic_bytes_per_solve": nbytes, "fine_unknowns": n, "cycles_per_solve": cycles,
            "updates_per_solve": updates, "pair_passes_per_solve": pair_passes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--ppc", type=int, default=0, help="override particles per cell (0 = config)")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="",
                    help="override the workload's particle precision (profiling the other kernel instance)")
    ap.add_argument("--injection", choices=["device", "host"], default="device",
                    help="boundary injection drawn on the device (Philox, same law) or on the host (mt19937)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cmatrix-cache", default="")
    ap.add_argument("--guess-weights", default=os.path.join(ROOT, "fopis_amd", "weights", "guess_c4_corr.npz"),
                    help="trained learned-guess weights for the Poisson comparison (skipped if absent)")
    ap.add_argument("--guess-steps", type=int, default=5)
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="print only the device-to-device copy rate of six 512 MiB FP32 arrays (GB/s)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (scale for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_ceiling:
        print(json.dumps({"hbm_copy_gbs": copy_ceiling_gbs()}), flush=True)
        return
    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from fopis_amd import lib as F
    from fopis_amd.host import Host

    wl = WORKLOADS[args.workload]
    log(F.device_info())
    t0 = time.perf_counter()
    cfg_path = os.path.join(ROOT, "fopis_amd", "configs", wl["json"])
    tmp_cfg = None
    if args.precision:
        import tempfile

        with open(cfg_path) as f:
            cfg = json.load(f)
        cfg.setdefault("Build", {})["precision"] = args.precision
        for o in cfg.get("Object", {}).values():
            o["file_name"] = os.path.join(os.path.dirname(cfg_path), o["file_name"])
        fd, tmp_cfg = tempfile.mkstemp(suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(cfg, f)
        cfg_path = tmp_cfg
    h = Host(cfg_path, device_loader=True, device_seed=20251202,
             ppc_override=args.ppc, build_cmatrix=True, initialize=True, cmatrix_cache=args.cmatrix_cache,
             device_injection=args.injection == "device")
    if tmp_cfg:
        os.unlink(tmp_cfg)
    ctx = h.ctx
    nparts = sum(ctx.num_particles(s) for s in range(h.S))
    log(f"setup {time.perf_counter() - t0:.1f} s, {nparts} particles")
    for i in range(args.warmup):
        h.main_loop_es()
    ctx.synchronize()
    log("warmup done")
    ctx.reset_timers()
    counts = []
    gen_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        h.main_loop_es()
        gen_ms += h.generate_ms()
        if (i + 1) % max(1, args.steps // 5) == 0:
            log(f"step {i + 1}/{args.steps}")
    ctx.synchronize()
    wall = time.perf_counter() - t0
    counts = [ctx.num_particles(s) for s in range(h.S)]
    n_end = sum(counts)
    part_ms, n_part = ctx.phase_time("particles")
    tile_ms, n_tile = ctx.phase_time("tile")
    bin_ms, _ = ctx.phase_time("bin")
    red_ms, _ = ctx.phase_time("reduce")
    poi_ms, n_poi = ctx.phase_time("poisson")
    ef_ms, _ = ctx.phase_time("efield")
    first_ms, _ = ctx.phase_time("poisson_first")
    stats = F.solver_stats(ctx)
    n_avg = 0.5 * (nparts + n_end)
    value = n_avg / (part_ms * 1e-3)
    bpp = wl["bytes_per_particle"][h.desc.precision]
    achieved = bpp * n_avg / (tile_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if h.desc.precision == 0 else "f32",
        "data": "synthetic (device Philox Maxwellian load; %s)"
                % ("device Philox boundary injection" if args.injection == "device" else "host mt19937 injection"),
        "config": {"workload": wl["name"], "particles": int(n_avg), "ppc": args.ppc or None,
                   "poisson_tol": 1e-8, "injection": args.injection,
                   "mg": "V(%d,%d) red-black SOR omega=%g in %s, FP64 defect correction"
                         % (h.desc.mg_pre, h.desc.mg_post, h.desc.mg_omega,
                            "FP32" if h.desc.mg_precision == 1 else "FP64")},
        "push_deposit_ms": part_ms,
        "tile_kernel_ms": tile_ms,
        "bin_ms": bin_ms,
        "reduce_ms": red_ms,
        "poisson_ms": poi_ms,
        "poisson_solves_per_step": n_poi / max(1, n_part),
        "poisson_cycles_per_solve": stats["total_cycles"] / max(1, stats["total_solves"]),
        "poisson_first_solve_ms": first_ms,
        "poisson_first_solve_cycles": stats["first_cycles"] / max(1, args.steps),
        "efield_ms": ef_ms,
        "poisson_roofline": poisson_roofline(h.desc, poi_ms,
                                             stats["total_cycles"] / max(1, stats["total_solves"]),
                                             stats["updates"] / max(1, stats["total_solves"]),
                                             stats["pair_passes"] / max(1, stats["total_solves"])),
        "host_generate_ms_per_step": gen_ms / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_tile (gather+Boris+move+absorb+count+deposit+bin), all species per step",
                     "algorithmic_bytes_per_particle": bpp},
    }
    tr = pmc_traffic(args.workload) if not args.precision else None
    if tr:
        # HBM bytes of one k_tile launch (one species) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
        line["roofline"]["traffic"] = tr.get("bytes_per_launch")
        line["roofline"]["algorithmic_bytes_per_launch"] = bpp * n_avg / h.S
    # the first solve of one more step, for the CPU Poisson baseline: its
    # right-hand side (the deposited charge, before the conductor correction)
    # and its warm start (the previous step's phi)
    poisson_case = None
    if not args.no_cpu_baseline:
        phi_prev = ctx.get_field(F.PHI)
        h.main_loop_es()
        rho_new = sum(ctx.get_field(F.RHO, s + 1) for s in range(h.S))
        poisson_case = (phi_prev, rho_new)
    h.close()
    if args.guess_steps > 0 and os.path.exists(args.guess_weights):
        log("learned-guess leg (C4)")
        try:
            line["poisson_learned_vs_plain"] = learned_leg(args.guess_steps, args.guess_weights)
        except Exception as e:  # reported, never silently replaced
            line["poisson_learned_vs_plain"] = {"error": str(e)}
    # torch ships its own HIP runtime, which cannot share this process with the
    # library's: the copy runs in a child process
    try:
        import subprocess

        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-ceiling"], capture_output=True,
                           text=True, timeout=300)
        line["hbm_copy_gbs"] = float(json.loads(r.stdout.strip().splitlines()[-1])["hbm_copy_gbs"])
        line["roofline"]["frac_of_copy"] = achieved / line["hbm_copy_gbs"]
    except Exception as e:  # reported, never silently replaced
        line["hbm_copy_gbs"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        info = host_info()
        threads = cpu_threads(info)
        log(f"cpu baseline on {threads} threads ({info})")
        try:
            line["cpu_baseline"] = cpu_baseline(args.workload, threads, info)
        except Exception as e:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": str(e)}
        log("cpu Poisson baseline")
        try:
            line["poisson_cpu_baseline"] = poisson_cpu_baseline(
                h.desc, args.workload, poisson_case[0], poisson_case[1], threads, first_ms,
                stats["first_cycles"] / max(1, args.steps))
        except Exception as e:  # reported, never silently replaced
            line["poisson_cpu_baseline"] = {"error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
