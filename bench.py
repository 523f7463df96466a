"""Benchmark: particle-steps/s of the PM gravity step on one MI355X.

Workload (default): the two-disk encounter, 2 x 2,097,152 particles, G=128
(S=256 zero-padded isolated FFT), dt = 1e-3, Morton re-sort every 10 steps.
Each step is the whole PM step (sort when due, FFT Poisson solve, gradient,
fused gather / kick / drift / deposit, reductions) with inputs resident in HBM.

Prints one JSON line. ``roofline`` covers the dominant kernel (k_integrate),
timed with HIP events on the library's stream inside the timed region;
``force_error`` gives the force rel-RMS against the direct sum (the metric's
second half) for the benchmarked step and for the learned-correction
comparison (plain G=128, G=128 + network, plain G=256) with the corrected
step's rates; ``cpu_baseline`` times the CPU oracle's reference-order step on
a bounded sample of the same workload.
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


def integrate_bytes(n_live, G, contact=False, deposit=True):
    """Algorithmic bytes of one k_integrate launch: pos + vel read and written
    (64 B per particle) + the potential read once (4 B per cell; the gather
    takes its differences) + the fixed-point density written once (8 B per
    cell) when the integrate deposits (+ the contact acceleration, 16 B per
    particle, when contact is on). Sparse contact steps deposit in a gather
    kernel of their own instead (deposit=False)."""
    return (80.0 if contact else 64.0) * n_live + (12.0 if deposit else 4.0) * G ** 3


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, n_steps):
    """The oracle's reference-order step (FP32 particles and FP32 mesh, OpenMP
    threads = OMP_NUM_THREADS, else every host core) on the same particles."""
    from oracle import oracle as O
    n = cfg["pos"].shape[0]
    prm = O.make_params(cfg["G"], gravity=cfg["gravity"], softeningCells=cfg["soft"],
                        sortInterval=cfg.get("sort_interval", 10), keyMode=O.KEY_MORTON,
                        fixedDt=cfg["dt"], allocN=n, fp32Mesh=1)
    st = O.Stepper(prm, cfg["pos"], cfg["vel"])
    st.step(1)  # builds and caches the Green spectrum, warms the allocator
    t0 = time.perf_counter()
    st.step(n_steps)
    dt = time.perf_counter() - t0
    return n * n_steps / dt, dt


def force_errors(cfg, sim, log_fn):
    """Force rel-RMS against the Plummer direct sum on 16,384 fixed particles
    (method of proto/cuda/x3.cu:109-185): the benchmarked step as run, and the
    learned-correction comparison at the initial positions (mesh softening 1
    cell): plain G=128, G=128 + the packaged network, plain G=256; plus the
    corrected step's rate refreshed every step, every 2 and every 5 steps
    against plain G=256."""
    from stardust_amd import force_error as FE
    from stardust_amd import learned
    eps = learned.EPS_BOX_DEFAULT
    out = {"targets": 16384, "direct_sum_eps_box": eps}
    r = FE.evaluate(sim, eps)
    out["benchmarked_step"] = {"G": cfg["G"], "softening_cells": cfg["soft"], "rel_rms": round(r["rel_rms"], 5),
                               "lsq_scale": round(r["lsq_scale"], 4), "at": "positions after the timed steps"}
    p4, v4, M, L = cfg["pos"], cfg["vel"], cfg["M"], cfg["L"]
    try:
        mlp, out_scale = learned.load_model()
    except OSError:
        mlp, out_scale = None, 1.0
        log_fn("no packaged network: learned-correction lines skipped")
    err = learned.evaluate(p4, v4, M, L, 128, 256, mlp, out_scale)
    out["softening_cells"] = 1.0
    for tag, e in err.items():
        out[tag.replace("+", "_")] = {"G": e["G"], "rel_rms": round(e["rel_rms"], 5), "lsq_scale": round(e["lsq_scale"], 4)}
    if mlp is not None:
        rate_f, ms_f = learned.time_steps(p4, v4, M, L, 256)
        rates = {"plain_G256": {"particle_steps_per_s": round(rate_f, 1), "ms_per_step": round(ms_f, 4)}}
        for k in (1, 2, 5):
            v, ms = learned.time_steps(p4, v4, M, L, 128, mlp, out_scale, interval=k)
            rates[f"G128_mlp_refresh{k}"] = {"particle_steps_per_s": round(v, 1), "ms_per_step": round(ms, 4),
                                             "vs_plain_G256": round(v / rate_f, 3)}
        out["rates"] = rates
    return out


def copy_calibration(mib):
    """Plain device-to-device copies of `mib` MiB (the runtime's copy kernel),
    so a counter run sees a kernel of known traffic beside k_integrate."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) or hip.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(a, 1, nbytes)
    for _ in range(3):
        if hip.hipMemcpy(b, a, nbytes, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in the copy calibration")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {mib} MiB device-to-device")


def _counter_rows(path, counter):
    import csv
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            key = (r["Dispatch_Id"], r["Kernel_Name"])
            out[key] = out.get(key, 0.0) + float(r["Counter_Value"])
    return out


TWO_DISK = "two-disk 4,194,304 particles, G=128"


def _kernel_counters(path, name, kernel, mib):
    rows = _counter_rows(path, name)
    sel = [v for (d, k), v in rows.items() if kernel in k]
    # the calibration copies are the last three copy-kernel dispatches
    copy = [v for (d, k), v in sorted(rows.items(), key=lambda kv: int(kv[0][0])) if "copyBuffer" in k][-3:]
    if not sel or len(copy) < 3:
        raise RuntimeError(f"{path}: no {kernel} or calibration copy dispatches for {name}")
    kib = sorted(sel)[len(sel) // 2]
    copy_kib = sorted(copy)[1]
    factor = (mib * 1024.0) / copy_kib
    return {"kernel_kib_median": kib, "copy_kib_median": copy_kib, "copy_kib_last3": copy, "scale": factor,
            "bytes": kib * 1024.0 * factor, "dispatches": len(sel)}


def traffic_summary(fetch_csv, write_csv, mib, out_path, command, kernels=("k_integrate",), workload=TWO_DISK):
    """HBM bytes per launch of each kernel (name substrings, the first one the
    roofline kernel) from two counter runs (FETCH_SIZE and WRITE_SIZE, KiB per
    dispatch, median over dispatches), each scaled by what the same counter
    reads for the copy kernel of `mib` MiB (reads mib, writes mib)."""
    per = {}
    for kern in kernels:
        res = {name: _kernel_counters(path, name, kern, mib)
               for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv))}
        per[kern] = {"counters": res,
                     "hbm_bytes_per_launch": round(res["FETCH_SIZE"]["bytes"] + res["WRITE_SIZE"]["bytes"])}
    first = per[kernels[0]]
    doc = {"kernel": "sd::" + kernels[0], "workload": workload, "copy_calibration_mib": mib,
           "counters": first["counters"], "hbm_bytes_per_launch": first["hbm_bytes_per_launch"],
           "command": command}
    if len(kernels) > 1:
        doc["other_kernels"] = {k: per[k] for k in kernels[1:]}
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc, indent=1))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic-summary":
        # offline: bench.py --traffic-summary FETCH.csv WRITE.csv MIB OUT.json "COMMAND"
        #          [KERNEL ...] [--workload TEXT]   (KERNEL: name substrings, default k_integrate)
        fetch_csv, write_csv, mib, out_path, command = sys.argv[2:7]
        rest = sys.argv[7:]
        workload = TWO_DISK
        if "--workload" in rest:
            i = rest.index("--workload")
            workload = rest[i + 1]
            rest = rest[:i] + rest[i + 2:]
        traffic_summary(fetch_csv, write_csv, int(mib), out_path, command, tuple(rest) or ("k_integrate",),
                        workload)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-force-error", action="store_true")
    ap.add_argument("--kernel", default="integrate", help="kernel timed for the roofline")
    ap.add_argument("--workload", choices=("two_disk", "cosmic_web"), default="two_disk",
                    help="two_disk: the headline 4M-particle G=128 step; cosmic_web: 16.7M particles, "
                         "periodic G=512 with contact")
    ap.add_argument("--timer-every", type=int, default=7,
                    help="bracket one launch in N of the roofline kernel with HIP events")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not provided",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import stardust_amd as sd
    from stardust_amd import ic

    log("building initial conditions")
    contact = args.workload == "cosmic_web"
    if contact:
        # 256^3 lattice, Zel'dovich displacements (P(k) ~ k^-2, rms 0.3 cells), v = 0,
        # periodic G = 512, contact radius 0.2 cells, re-sort every step, CFL dt
        pos, vel = ic.zeldovich(256, seed=1, rms_cells=0.3, G_mesh=512)
        cfg = {"pos": pos, "vel": vel, "G": 512}
        n, G = pos.shape[0], 512
        conf = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC,
                                 gravity=ic.gravity_periodic(G, 1.0, 1.0), softeningCells=3.0, sortInterval=1,
                                 contactEnabled=1, contactRadiusCells=0.2)
        sim = sd.Sim(conf, sd.default_modes())
        sim.set_particles(pos, vel)
        desc = ("cosmic web: 256^3 lattice (16,777,216 particles) with Zel'dovich displacements "
                "(P(k)~k^-2, rms 0.3 cells, v=0), periodic G=512, contact radius 0.2 cells, "
                "re-sort every step, CFL dt")
        metric = "particle-steps/s (16.7M, periodic 512^3 PM + contact, 1 GPU)"
        fft_side = G
    else:
        cfg = ic.config1(G=args.grid)
        n = cfg["pos"].shape[0]
        G = cfg["G"]
        conf = sd.default_config(particleCount=n, gridSize=G, gravity=cfg["gravity"],
                                 softeningCells=cfg["soft"], sortInterval=cfg["sort_interval"])
        sim = sd.Sim(conf, sd.default_modes())
        sim.set_fixed_dt(cfg["dt"])
        sim.set_particles(cfg["pos"], cfg["vel"])
        desc = (f"two-disk encounter, 2x2,097,152 particles, G={G} / S={2 * G} isolated PM, dt=1e-3, "
                "Morton re-sort every 10 steps")
        metric = "particle-steps/s (4M, 256^3 PM, 1 GPU)"
        fft_side = 2 * G
    log(f"N={n} G={G} warmup={args.warmup}")
    sim.run(args.warmup)
    sim.synchronize()
    # one launch in 7 bracketed by events (each bracket adds two queue markers; 7 is
    # coprime with the sort interval, so sort steps are sampled at their share)
    sim.kernel_timer(args.kernel, every=args.timer_every)
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
        # the cosmic web's sparse contact steps (0.125 particles per cell) deposit
        # in the gather kernel, not in the integrate
        gather = contact and n < 0.5 * G ** 3
        nbytes = integrate_bytes(n_live, G, contact, deposit=not gather)
        ach = nbytes / avg_s / 1e9
        # HBM traffic needs counter runs (rocprofv3 --pmc), which this run does
        # not make: `traffic` stays null and the committed counter summary of
        # the same kernel and workload is quoted beside it, with its source
        profiled = None
        tname = ("traffic_cosmic_web_16M_G512.json" if contact and G == 512
                 else "traffic_k_integrate_two_disk_4M_G128.json" if not contact and G == 128 else None)
        tpath = os.path.join(ROOT, "profiles", tname) if tname else None
        if args.kernel == "integrate" and tpath and os.path.exists(tpath):
            with open(tpath) as f:
                profiled = {"bytes_per_launch": json.load(f).get("hbm_bytes_per_launch"),
                            "source": f"profiles/{tname} (FETCH_SIZE and WRITE_SIZE counter runs of bench.py, "
                                      "not of this run)"}
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "traffic_profiled": profiled,
                "kernel": "sd::k_integrate", "avg_launch_us": round(avg_s * 1e6, 3),
                "algorithmic_bytes_per_launch": nbytes}

    ferr = None
    if not args.no_force_error and not contact and G == 128:
        log("force errors against the direct sum (16,384 particles)")
        ferr = force_errors(cfg, sim, log)

    cpu = None
    if not args.no_cpu and not contact:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        log(f"cpu baseline: oracle reference step, FP32 mesh, {args.cpu_steps} steps, {threads} threads")
        v, secs = cpu_baseline(cfg, args.cpu_steps)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": f"two-disk 4,194,304 particles, G={G}, {args.cpu_steps} steps after one warm-up step "
                         f"({secs:.1f} s): reference kernel order, FP32 particles and FP32 FFT mesh, OpenMP"}

    out = {
        "metric": metric,
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
        "config": {"workload": desc, "particles": n, "mesh": G, "fft_side": fft_side,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "force_error": ferr,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sim.close()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)


if __name__ == "__main__":
    main()
