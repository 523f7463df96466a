"""Batched ensemble: many independent disk scenes of uneven particle count and
mesh size stepped together (one launch per kernel for all scenes, a whole run
recorded into one HIP graph and launched once), with position snapshots that
feed the learned correction's training pairs.

Run ``python -m stardust_amd.batch --help`` for the benchmark report: the
64-scene ensemble (N_i log-uniform in [32,768, 1,048,576], G_i in {64, 96,
128}, 200 steps, snapshots every 20) timed as one graph launch, next to the
same scenes stepped one after another through the single-scene entry point.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import Batch, Sim, default_config, default_modes
from . import force_error as FE
from . import ic


def scene_configs(scenes, sort_interval=10):
    """Library configurations of (n, G, seed) scenes; returns (configs, data)."""
    configs, data = [], []
    for n, G, seed in scenes:
        d = ic.disk_scene(n, G, seed)
        configs.append(default_config(particleCount=n, gridSize=G, gravity=d["gravity"],
                                      softeningCells=d["soft"], sortInterval=sort_interval))
        data.append(d)
    return configs, data


def make_batch(scenes, sort_interval=10, log=None):
    """A Batch holding the scenes' initial conditions (fixed dt per scene)."""
    configs, data = scene_configs(scenes, sort_interval)
    b = Batch(configs)
    for i, d in enumerate(data):
        sc = b.scene(i)
        sc.set_particles(d["pos"], d["vel"])
        sc.set_fixed_dt(d["dt"])
    if log:
        log(f"[batch] {len(scenes)} scenes, {b.total_particles} particles")
    return b, data


def run_sequential(scenes, data, steps, sort_interval=10):
    """Seconds to step every scene alone (single-scene entry point, exact mesh
    sizes), one scene after another; allocation and upload excluded."""
    total = 0.0
    for (n, G, seed), d in zip(scenes, data):
        s = Sim(default_config(particleCount=n, gridSize=G, gravity=d["gravity"], softeningCells=d["soft"],
                               sortInterval=sort_interval), default_modes(exactGrid=1))
        s.set_particles(d["pos"], d["vel"])
        s.set_fixed_dt(d["dt"])
        s.synchronize()
        t0 = time.perf_counter()
        s.run(steps)
        s.synchronize()
        total += time.perf_counter() - t0
        s.close()
    return total


def pairs_from_snapshots(batch, data, G_train=128, n_targets=65536, scenes=None, snaps=None,
                         eps_box=1.0 / 256.0, soft=1.0, log=print):
    """Training pairs for the learned correction from batch snapshots: each
    chosen (scene, snapshot) is loaded on the G_train mesh (softening ``soft``
    cells) and gives, for up to ``n_targets`` particles, the 30 features and
    the residual (direct sum - PM) / a_unit. Returns X [m x 30], T [m x 3]."""
    nsnap = batch.snapshot_count()
    scenes = range(len(data)) if scenes is None else scenes
    snaps = range(nsnap) if snaps is None else snaps
    Xs, Ts = [], []
    for i in scenes:
        d = data[i]
        n = d["pos"].shape[0]
        gravity = ic.gravity_isolated(G_train, d["M"], d["L"])
        s = Sim(default_config(particleCount=n, gridSize=G_train, gravity=gravity, softeningCells=soft),
                default_modes())
        zero = np.zeros((n, 4), np.float32)
        for k in snaps:
            pos = batch.snapshot(k, i)
            s.set_particles(pos, zero)
            tg = FE.pick_targets(pos, min(n_targets, n), seed=1000 * i + k)
            X = s.features(tg)
            a_dir = s.direct_sum(tg, FE.direct_gm(gravity, G_train, n), eps_box)
            a_pm = s.particle_accel()[tg, :3].astype(np.float64)
            a_unit = gravity * G_train / n
            Xs.append(X)
            Ts.append(((a_dir - a_pm) / a_unit).astype(np.float32))
        s.close()
        if log:
            log(f"[batch] pairs from scene {i} (N={n}): {sum(x.shape[0] for x in Xs)} so far")
    return np.concatenate(Xs), np.concatenate(Ts)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--snapshot-every", type=int, default=20)
    ap.add_argument("--sort-interval", type=int, default=10)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--max-n", type=int, default=2 ** 20)
    ap.add_argument("--sequential", action="store_true", help="also time the scenes stepped one by one")
    ap.add_argument("--json", default=None, help="write the report here")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    scenes = ic.config3(args.scenes, seed=args.seed, n_range=(2 ** 15, args.max_n))
    t0 = time.perf_counter()
    b, data = make_batch(scenes, args.sort_interval, log)
    log(f"[batch] initial conditions + upload {time.perf_counter() - t0:.1f} s")
    N = b.total_particles
    t0 = time.perf_counter()
    b.run(args.steps, args.snapshot_every)
    b.synchronize()
    wall_first = time.perf_counter() - t0
    dev_ms = b.last_run_ms()
    nodes = b.graph_nodes()
    log(f"[batch] first run (record + launch) {wall_first:.3f} s wall, graph {nodes} nodes, "
        f"launch {dev_ms:.1f} ms device")
    # a second run replays the recorded graph
    t0 = time.perf_counter()
    b.run(args.steps, args.snapshot_every)
    b.synchronize()
    wall_replay = time.perf_counter() - t0
    dev_ms2 = b.last_run_ms()
    masses = [b.scene(i).measure_total_grid_mass() for i in range(len(scenes))]
    mass_ok = all(abs(m - n) <= 1e-6 * n for m, (n, _, _) in zip(masses, scenes))
    rep = {
        "workload": f"{len(scenes)} disks, N_i log-uniform in [32768, {args.max_n}], G_i in {{64, 96, 128}} "
                    f"(S = 2G, isolated), dt 1e-3, {args.steps} steps, snapshots every {args.snapshot_every}, "
                    f"sort every {args.sort_interval}",
        "scenes": len(scenes),
        "particles": N,
        "meshes": {str(G): sum(1 for s in scenes if s[1] == G) for G in (64, 96, 128)},
        "graph_nodes": nodes,
        "record_and_launch_s": round(wall_first, 4),
        "launch_ms": round(dev_ms, 3),
        "replay_ms": round(dev_ms2, 3),
        "replay_wall_s": round(wall_replay, 4),
        "particle_steps_per_s": round(N * args.steps / (dev_ms2 / 1e3), 1),
        "mass_conserved": mass_ok,
    }
    if args.sequential:
        # the state after 2 runs is irrelevant here: fresh scenes, same steps
        secs = run_sequential(scenes, data, args.steps, args.sort_interval)
        rep["sequential_s"] = round(secs, 4)
        rep["sequential_particle_steps_per_s"] = round(N * args.steps / secs, 1)
        rep["batch_speedup"] = round(secs / (dev_ms2 / 1e3), 3)
    b.close()
    print(json.dumps(rep), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
