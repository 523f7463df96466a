"""Benchmark: particle-steps/s of the PM gravity step on one MI355X.

Workload (default): the two-disk encounter, 2 x 2,097,152 particles, G=128
(S=256 zero-padded isolated FFT), dt = 1e-3, Morton re-sort every 10 steps.
Each step is the whole PM step (sort when due, FFT Poisson solve, gradient,
fused gather / kick / drift / deposit, reductions) with inputs resident in HBM.

Prints one JSON line. ``roofline`` covers the dominant kernel (k_integrate),
timed with HIP events on the library's stream inside the timed region;
``cpu_baseline`` times the CPU oracle's reference-order step on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def integrate_bytes(n_live, G):
    """Algorithmic bytes of one k_integrate launch: pos + vel read and written
    (64 B per particle) + the acceleration grid read once (16 B per cell) + the
    fixed-point density written once (8 B per cell)."""
    return 64.0 * n_live + 24.0 * G ** 3


def cpu_baseline(cfg, n_steps, threads):
    from oracle import oracle as O
    import numpy as np
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    n = cfg["pos"].shape[0]
    prm = O.make_params(cfg["G"], gravity=cfg["gravity"], softeningCells=cfg["soft"],
                        sortInterval=cfg.get("sort_interval", 10), keyMode=O.KEY_MORTON,
                        fixedDt=cfg["dt"], allocN=n)
    st = O.Stepper(prm, cfg["pos"], cfg["vel"])
    st.step(1)  # builds and caches the Green spectrum, warms the allocator
    t0 = time.perf_counter()
    st.step(n_steps)
    dt = time.perf_counter() - t0
    return n * n_steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default="integrate", help="kernel timed for the roofline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not provided",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import stardust_amd as sd
    from stardust_amd import ic

    log("building initial conditions")
    cfg = ic.config1(G=args.grid)
    n = cfg["pos"].shape[0]
    G = cfg["G"]
    conf = sd.default_config(particleCount=n, gridSize=G, gravity=cfg["gravity"],
                             softeningCells=cfg["soft"], sortInterval=cfg["sort_interval"])
    sim = sd.Sim(conf, sd.default_modes())
    sim.set_fixed_dt(cfg["dt"])
    sim.set_particles(cfg["pos"], cfg["vel"])
    log(f"N={n} G={G} warmup={args.warmup}")
    sim.run(args.warmup)
    sim.synchronize()
    sim.kernel_timer(args.kernel)
    log("timed region")
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    kms, launches = sim.kernel_time_ms()
    sim.kernel_timer(None)
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    n_live = sim.active_count()
    mass = sim.measure_total_grid_mass()
    log(f"{ms_per_step:.4f} ms/step, {value:.4e} particle-steps/s, mass {mass}")

    roof = None
    if launches > 0:
        avg_s = kms / launches / 1e3
        nbytes = integrate_bytes(n_live, G)
        ach = nbytes / avg_s / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic_k_integrate_two_disk_4M_G128.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "sd::k_integrate", "avg_launch_us": round(avg_s * 1e6, 3),
                "algorithmic_bytes_per_launch": nbytes}

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle reference step, {args.cpu_steps} steps, {threads} threads")
        v, secs = cpu_baseline(cfg, args.cpu_steps, threads)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"two-disk 4,194,304 particles, G={G}, {args.cpu_steps} steps after one "
                         f"warm-up step ({secs:.1f} s): FP64 FFT mesh + FP32 particle kernels"}

    out = {
        "metric": "particle-steps/s (4M, 256^3 PM, 1 GPU)",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "two-disk encounter, 2x2,097,152 particles, G=128 / S=256 "
                               "isolated PM, dt=1e-3, Morton re-sort every 10 steps",
                   "particles": n, "mesh": G, "fft_side": 2 * G, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sim.close()


if __name__ == "__main__":
    main()
