"""Learned sub-cell correction: training data, training and evaluation.

Pipeline (all heavy work on the GPU through the C ABI):
  1. scenes: exponential disks (and two-disk encounters) of varying particle
     count and box size, evolved a few PM steps; snapshots are taken at the
     mesh being corrected (G, softening 1 cell);
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


def train(X, T, epochs=30, batch=8192, lr=1e-3, seed=1, log=print):
    out_scale = float(np.sqrt((T.astype(np.float64) ** 2).mean()))
    m = Mlp(seed=seed)
    t0 = time.perf_counter()
    losses = m.train(X, T / out_scale, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"[learned] trained {epochs} epochs on {X.shape[0]} pairs in {time.perf_counter() - t0:.1f} s; "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}; outScale {out_scale:.4f}")
    return m, out_scale, losses


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
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    X, T = training_pairs(G=args.coarse, n_targets=args.targets, log=log)
    mlp, out_scale, losses = train(X, T, epochs=args.epochs, log=log)
    p4, v4, M, L = scene_pair(args.eval_each, seed=1)
    err = evaluate(p4, v4, M, L, args.coarse, args.fine, mlp, out_scale)
    log("[learned] errors", json.dumps(err))
    plain_f, ms_f = time_steps(p4, v4, M, L, args.fine)
    corr_c, ms_c = time_steps(p4, v4, M, L, args.coarse, mlp, out_scale)
    corr_c5, ms_c5 = time_steps(p4, v4, M, L, args.coarse, mlp, out_scale, interval=5)
    plain_c, ms_pc = time_steps(p4, v4, M, L, args.coarse)
    rep = {"errors": err, "train_pairs": int(X.shape[0]), "epochs": args.epochs,
           "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "out_scale": out_scale,
           "eps_box": EPS_BOX_DEFAULT, "eval_particles": int(p4.shape[0]),
           "particle_steps_per_s": {"plain_fine": plain_f, "coarse_mlp": corr_c,
                                    "coarse_mlp_every5": corr_c5, "plain_coarse": plain_c},
           "ms_per_step": {"plain_fine": ms_f, "coarse_mlp": ms_c, "coarse_mlp_every5": ms_c5,
                           "plain_coarse": ms_pc},
           "speedup_coarse_mlp_every5_vs_plain_fine": corr_c5 / plain_f,
           "speedup_coarse_mlp_vs_plain_fine": corr_c / plain_f,
           "coarse_mlp_error_vs_plain_fine": err["coarse+mlp"]["rel_rms"] / err["fine"]["rel_rms"]}
    print(json.dumps(rep), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(rep, f, indent=1)
    if args.save_model:
        W, b = mlp.get_params()
        np.savez(args.save_model, W=W, b=b, out_scale=out_scale)


if __name__ == "__main__":
    main()
