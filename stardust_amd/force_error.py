"""Force-error metric of the PM step against a direct sum.

Method of the reference's prototype proto/cuda/x3.cu:109-185 (there in 2D):
the PM acceleration at each target particle is compared with a softened
direct sum over all particles,

    rel_rms = sqrt(sum |a_pm - a_dir|^2) / sqrt(sum |a_dir|^2),

and the least-squares scale a = <a_pm, a_dir> / <a_pm, a_pm> (x3.cu:168-177) is
reported as a diagnostic. Here the direct sum is Plummer-softened with one
fixed length at every mesh resolution, in box units, and both sides run on the
GPU (sd_particle_accel, sd_direct_sum).

Units: on the isolated mesh G_N,box * m = gravity / (G * N) per particle
(pot = (gravity/N) * sum -1/(r_cells + soft), r_cells = G r_box).
"""
from __future__ import annotations

import numpy as np


def direct_gm(gravity: float, G: int, n: int, periodic: bool = False) -> float:
    """G_N,box times the particle mass implied by the mesh units."""
    if periodic:
        return gravity / (4.0 * np.pi * G ** 3 * n)
    return gravity / (G * n)


def pick_targets(pos: np.ndarray, count: int = 16384, seed: int = 12345) -> np.ndarray:
    live = np.nonzero(pos[:, 0] >= 0)[0]
    rng = np.random.default_rng(seed)
    count = min(count, live.shape[0])
    return np.sort(rng.choice(live, count, replace=False)).astype(np.int32)


def rel_rms(approx: np.ndarray, truth: np.ndarray):
    a = np.asarray(approx, np.float64).reshape(-1, 3)
    t = np.asarray(truth, np.float64).reshape(-1, 3)
    se = ((a - t) ** 2).sum()
    st = (t ** 2).sum()
    aa = (a ** 2).sum()
    scale = float((a * t).sum() / aa) if aa > 0 else 0.0
    return float(np.sqrt(se / st)) if st > 0 else 0.0, scale


def evaluate(sim, eps_box: float, targets: np.ndarray | None = None, count: int = 16384,
             seed: int = 12345):
    """PM-vs-direct error of the simulation's current positions.

    Returns dict(rel_rms, lsq_scale, a_pm, a_dir, targets)."""
    cfg = sim.config()
    pos, _ = sim.get_particles()
    if targets is None:
        targets = pick_targets(pos, count, seed)
    n = sim.particle_count()
    gm = direct_gm(cfg.gravity, sim.grid_size(), n, cfg.boundary == 1)
    a_dir = sim.direct_sum(targets, gm, eps_box)
    a_pm = sim.particle_accel()[targets, :3].astype(np.float64)
    err, scale = rel_rms(a_pm, a_dir)
    return dict(rel_rms=err, lsq_scale=scale, a_pm=a_pm, a_dir=a_dir, targets=targets)
