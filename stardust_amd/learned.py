"""Learned sub-cell correction: training data, training and evaluation.

Pipeline (all heavy work on the GPU through the C ABI):
  1. scenes: the batched ensemble (64 disks of 2^15..2^20 particles advanced in
     one graph launch, snapshots every 20 steps) plus exponential disks and
     two-disk encounters of 2^18..2^22 particles evolved a few PM steps; each
     snapshot is re-meshed at the mesh being corrected (G, softening 1 cell);
  2. pairs: for a subsample of particles per snapshot, the 30 features
     (sd_features), the PM acceleration (sd_particle_accel) and the Plummer
     direct sum over all particles (sd_direct_sum, fixed softening eps_box);
     the target is (a_direct - a_pm) / a_unit, a_unit = gravity * G / N;
  3. training: Adam on the BF16 MFMA network (sd_mlp_train); targets are
     divided by their RMS (outScale) so the network sees unit variance;
  4. evaluation: rel-RMS force error versus direct sum on a fixed subsample of
     the evaluation scene, for the corrected coarse mesh and plain meshes.

Run ``python -m stardust_amd.learned --help`` for the command-line report.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import Mlp, Sim, default_config, default_modes
from . import force_error as FE
from . import ic

EPS_BOX_DEFAULT = 1.0 / 256.0  # one cell of the 256-mesh (S = 512), every resolution


def _sim_for(pos, vel, G, gravity, soft=1.0, sort_interval=10):
    n = pos.shape[0]
    s = Sim(default_config(particleCount=n, gridSize=G, gravity=gravity, softeningCells=soft,
                           sortInterval=sort_interval), default_modes())
    s.set_particles(pos, vel)
    return s


def scene_disk(n, seed, L=None):
    """Single exponential disk scaled into a box of side L."""
    pos, vel = ic.exp_disk(n, seed=seed)
    L = L or 12.0
    p4, v4 = ic.to_box(pos, vel, (0.0, 0.0, 0.0), L)
    return p4, v4, 1.0, L


def scene_pair(n_each, seed, L=20.0):
    pos, vel, M = ic.two_disks(n_each, seed=seed)
    p4, v4 = ic.to_box(pos, vel, (4.0, 0.3, 0.0), L)
    return p4, v4, M, L


def training_pairs(G=128, scenes=None, n_targets=65536, eps_box=EPS_BOX_DEFAULT, evolve=(0, 20),
                   dt=1e-3, log=print):
    """Returns X [m x 30], T [m x 3] (residual / a_unit)."""
    if scenes is None:
        rng = np.random.default_rng(2024)
        scenes = []
        for k in range(8):
            n = int(round(math.exp(rng.uniform(math.log(2 ** 18), math.log(2 ** 22)))))
            # boxes wide enough for the scene (disk radius 5; the pair spans x in [-5, 13])
            L = float(rng.uniform(20.0, 26.0) if k % 2 else rng.uniform(11.0, 24.0))
            if k % 2:
                scenes.append(("pair", n // 2, 100 + k, L))
            else:
                scenes.append(("disk", n, 100 + k, L))
    Xs, Ts = [], []
    for kind, n, seed, L in scenes:
        if kind == "disk":
            p4, v4, M, L = scene_disk(n, seed, L)
        else:
            p4, v4, M, L = scene_pair(n, seed, L)
        N = p4.shape[0]
        gravity = ic.gravity_isolated(G, M, L)
        s = _sim_for(p4, v4, G, gravity)
        s.set_fixed_dt(dt)
        done = 0
        for steps in evolve:
            if steps > done:
                s.run(steps - done)
                done = steps
            pos, _ = s.get_particles()
            tg = FE.pick_targets(pos, n_targets, seed=seed * 7 + steps)
            X = s.features(tg)
            gm = FE.direct_gm(gravity, G, N)
            a_dir = s.direct_sum(tg, gm, eps_box)
            a_pm = s.particle_accel()[tg, :3].astype(np.float64)
            a_unit = gravity * G / N
            Xs.append(X)
            Ts.append(((a_dir - a_pm) / a_unit).astype(np.float32))
            log(f"[learned] scene {kind} N={N} L={L:.1f} step {steps}: {tg.size} pairs, "
                f"residual rms {np.sqrt((Ts[-1] ** 2).mean()):.4f}")
        s.close()
    return np.concatenate(Xs), np.concatenate(Ts)


def training_pairs_batch(G=128, targets_per_snapshot=2048, n_scenes=64, steps=200, every=20, seed=3,
                         log=print):
    """Pairs from the batched ensemble: 64 config-0 disks of 2^15..2^20
    particles on meshes {64, 96, 128}, advanced ``steps`` steps in one graph
    launch with position snapshots every ``every`` steps; each snapshot is
    re-meshed at G (softening 1 cell) for features, PM force and direct sum."""
    from . import batch as BT
    scenes = ic.config3(n_scenes, seed=seed)
    b, data = BT.make_batch(scenes, 10, log)
    b.run(steps, every)
    b.synchronize()
    X, T = BT.pairs_from_snapshots(b, data, G_train=G, n_targets=targets_per_snapshot, log=None)
    b.close()
    log(f"[learned] {X.shape[0]} pairs from {len(scenes)} batched scenes x {steps // every} snapshots, "
        f"residual rms {np.sqrt((T.astype(np.float64) ** 2).mean()):.2f}")
    return X, T


def train(X, T, epochs=30, batch=8192, lr=1e-3, seed=1, log=print):
    out_scale = float(np.sqrt((T.astype(np.float64) ** 2).mean()))
    m = Mlp(seed=seed)
    t0 = time.perf_counter()
    losses = m.train(X, T / out_scale, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"[learned] trained {epochs} epochs on {X.shape[0]} pairs in {time.perf_counter() - t0:.1f} s; "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}; outScale {out_scale:.4f}")
    return m, out_scale, losses


MODEL_TWO_DISK_G128 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "mlp_two_disk_G128.npz")


def load_model(path=MODEL_TWO_DISK_G128):
    """A network written by ``--save-model`` (arrays only, no pickles):
    returns (Mlp, out_scale). The packaged model was trained for the G=128
    mesh by ``python -m stardust_amd.learned --save-model <path>`` (default
    options: 2,359,296 pairs, 30 epochs)."""
    with np.load(path, allow_pickle=False) as d:
        W, b, out_scale = d["W"], d["b"], float(d["out_scale"])
    m = Mlp(seed=1)
    m.set_params(W, b)
    return m, out_scale


def evaluate(pos, vel, M, L, G_coarse, G_fine, mlp=None, out_scale=1.0, eps_box=EPS_BOX_DEFAULT,
             n_targets=16384, seed=12345):
    """Force errors of plain coarse, corrected coarse and plain fine meshes."""
    res = {}
    targets = FE.pick_targets(pos, n_targets, seed)
    a_dir = None
    for tag, G, use in (("coarse", G_coarse, False), ("coarse+mlp", G_coarse, True), ("fine", G_fine, False)):
        if use and mlp is None:
            continue
        gravity = ic.gravity_isolated(G, M, L)
        s = _sim_for(pos, vel, G, gravity)
        if a_dir is None:
            a_dir = s.direct_sum(targets, FE.direct_gm(gravity, G, pos.shape[0]), eps_box)
        a = s.particle_accel()[targets, :3].astype(np.float64)
        if use:
            s.attach_mlp(mlp, out_scale)
            a = a + s.mlp_correction()[targets, :3]
            s.attach_mlp(None)
        err, scale = FE.rel_rms(a, a_dir)
        res[tag] = {"G": G, "rel_rms": err, "lsq_scale": scale}
        s.close()
    return res


def evaluate_staleness(pos, vel, M, L, G_coarse, G_fine, mlp, out_scale, lags=9, dt=1e-3,
                       eps_box=EPS_BOX_DEFAULT, n_targets=16384, seed=12345):
    """Force error of a correction computed ``lag`` steps earlier. The
    corrected coarse run keeps its first correction (no re-sort, so particle
    order is fixed); at each lag the error of PM + that stale correction is
    compared with PM + a fresh correction, plain coarse and plain fine PM, all
    against the direct sum at the current positions."""
    n = pos.shape[0]
    grav_c = ic.gravity_isolated(G_coarse, M, L)
    grav_f = ic.gravity_isolated(G_fine, M, L)
    run = _sim_for(pos, vel, G_coarse, grav_c, sort_interval=0)
    run.set_fixed_dt(dt)
    run.attach_mlp(mlp, out_scale)
    run.set_mlp_interval(1 << 30)
    stale = run.mlp_correction()
    probe_c = _sim_for(pos, vel, G_coarse, grav_c, sort_interval=0)
    probe_c.attach_mlp(mlp, out_scale)
    probe_f = _sim_for(pos, vel, G_fine, grav_f, sort_interval=0)
    targets = FE.pick_targets(pos, n_targets, seed)
    gm = FE.direct_gm(grav_c, G_coarse, n)
    out = []
    for lag in range(lags + 1):
        if lag:
            run.run(1)
        p, v = run.get_particles()
        probe_c.set_particles(p, v)
        probe_f.set_particles(p, v)
        a_dir = probe_c.direct_sum(targets, gm, eps_box)
        a_c = probe_c.particle_accel()[targets, :3].astype(np.float64)
        fresh = probe_c.mlp_correction()[targets, :3]
        a_f = probe_f.particle_accel()[targets, :3].astype(np.float64)
        out.append({"lag": lag,
                    "coarse": FE.rel_rms(a_c, a_dir)[0],
                    "coarse+mlp_fresh": FE.rel_rms(a_c + fresh, a_dir)[0],
                    "coarse+mlp_stale": FE.rel_rms(a_c + stale[targets, :3], a_dir)[0],
                    "fine": FE.rel_rms(a_f, a_dir)[0]})
    probe_c.attach_mlp(None)
    run.attach_mlp(None)
    for s in (run, probe_c, probe_f):
        s.close()
    return out


def time_steps(pos, vel, M, L, G, mlp=None, out_scale=1.0, steps=20, warmup=5, dt=1e-3, interval=1):
    gravity = ic.gravity_isolated(G, M, L)
    s = _sim_for(pos, vel, G, gravity)
    s.set_fixed_dt(dt)
    if mlp is not None:
        s.attach_mlp(mlp, out_scale)
        s.set_mlp_interval(interval)
    s.run(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    if mlp is not None:
        s.attach_mlp(None)
    s.close()
    return pos.shape[0] * steps / wall, wall * 1e3 / steps


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--eval-each", type=int, default=2097152, help="particles per disk of the evaluation pair")
    ap.add_argument("--coarse", type=int, default=128)
    ap.add_argument("--fine", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--targets", type=int, default=65536)
    ap.add_argument("--out", default=None, help="write the JSON report here")
    ap.add_argument("--save-model", default=None, help="write parameters (.npz)")
    ap.add_argument("--pairs", choices=("batch", "scenes", "both"), default="both",
                    help="training pairs from the batched ensemble's snapshots (64 disks of 2^15..2^20 "
                         "particles), from 8 larger scenes (2^18..2^22 particles, disks and pairs), or both; "
                         "the ensemble alone does not reach the density range of the 4M-particle evaluation")
    ap.add_argument("--lags", type=int, default=9, help="staleness report up to this many steps")
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    parts = []
    if args.pairs in ("batch", "both"):
        parts.append(training_pairs_batch(G=args.coarse, log=log))
    if args.pairs in ("scenes", "both"):
        parts.append(training_pairs(G=args.coarse, n_targets=args.targets, log=log))
    X = np.concatenate([x for x, _ in parts])
    T = np.concatenate([t for _, t in parts])
    mlp, out_scale, losses = train(X, T, epochs=args.epochs, log=log)
    p4, v4, M, L = scene_pair(args.eval_each, seed=1)
    err = evaluate(p4, v4, M, L, args.coarse, args.fine, mlp, out_scale)
    log("[learned] errors", json.dumps(err))
    plain_f, ms_f = time_steps(p4, v4, M, L, args.fine)
    corr_c, ms_c = time_steps(p4, v4, M, L, args.coarse, mlp, out_scale)
    corr_c5, ms_c5 = time_steps(p4, v4, M, L, args.coarse, mlp, out_scale, interval=5)
    plain_c, ms_pc = time_steps(p4, v4, M, L, args.coarse)
    by_interval = {}
    for k in (2, 3, 5, 10):
        v, ms = time_steps(p4, v4, M, L, args.coarse, mlp, out_scale, interval=k)
        by_interval[str(k)] = {"particle_steps_per_s": v, "ms_per_step": ms, "speedup_vs_plain_fine": v / plain_f}
    stale = evaluate_staleness(p4, v4, M, L, args.coarse, args.fine, mlp, out_scale, lags=args.lags)
    log("[learned] staleness", json.dumps(stale))
    rep = {"errors": err, "train_pairs": int(X.shape[0]), "epochs": args.epochs,
           "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "out_scale": out_scale,
           "eps_box": EPS_BOX_DEFAULT, "eval_particles": int(p4.shape[0]),
           "particle_steps_per_s": {"plain_fine": plain_f, "coarse_mlp": corr_c,
                                    "coarse_mlp_every5": corr_c5, "plain_coarse": plain_c},
           "ms_per_step": {"plain_fine": ms_f, "coarse_mlp": ms_c, "coarse_mlp_every5": ms_c5,
                           "plain_coarse": ms_pc},
           "speedup_coarse_mlp_every5_vs_plain_fine": corr_c5 / plain_f,
           "speedup_coarse_mlp_vs_plain_fine": corr_c / plain_f,
           "coarse_mlp_error_vs_plain_fine": err["coarse+mlp"]["rel_rms"] / err["fine"]["rel_rms"],
           "pairs_source": args.pairs, "refresh_interval": by_interval, "staleness": stale}
    print(json.dumps(rep), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(rep, f, indent=1)
    if args.save_model:
        W, b = mlp.get_params()
        np.savez(args.save_model, W=W, b=b, out_scale=out_scale)


if __name__ == "__main__":
    main()
