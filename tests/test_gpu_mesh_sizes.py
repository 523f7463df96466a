"""Mesh parity at the FFT sizes the benchmark configurations run.

The transforms are planned per size (radix 3*8*8 at S=192, with a
register-resident 16*12 fused z pass; 16*16 with pruned first / last stages at
S=256; 8*8*8 with persistent prefetching column passes at S=512), so each size
is checked on its own:

* isolated G=96 (S=192) and G=128 (S=256) against the oracle's FP64 solve
  (``test_gpu_parity.test_potential_and_accel``);
* isolated G=96, G=128 and G=256 (S=512) against a closed form: point masses at
  cell centres deposit their whole count on one node each, and the isolated
  potential is then exactly sum_k -m_k / (r_k + soft) at every node (no
  periodic images, the zero padding removes them);
* periodic G=128, 256 and 512 against a closed form: a density that depends
  only on u = (x + b y + c z) mod G is a sum of plane waves along (1, b, c),
  so its potential is g(u), where g is the 1-D inverse DFT of the density's
  1-D spectrum times the reference multiplier -1/(4 pi^2 k^2) exp(-k^2 s^2/2)
  (Sim.cu:528-557), k the folded 3-D wavevector j (1, b, c) mod G. The
  gradient then has the closed form g(u + d) - g(u - d) per axis
  (Sim.cu:589-611).

Tolerances: the FP32 transforms against the FP64 forms, relative to the
largest potential value (2e-5, as at the smaller sizes); accelerations within
what that allows a central difference, 2 * 2e-5 * max|pot| * (G/2) potScale.
"""
import sys

import numpy as np
import pytest

import stardust_amd as sd
from closed_forms import plane_particles, plane_potential, point_mass_potential, point_masses

pytestmark = pytest.mark.gpu

POT_TOL = 2e-5


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("G", [96, 128, 256])
def test_isolated_point_masses_closed_form(gpu, oracle, G):
    _say("iso point masses", G)
    soft = 1.5
    gravity = 0.7
    src = point_masses(G)
    pos = []
    for x, y, z, m in src:
        for _ in range(m):
            pos.append(((x + 0.5) / G, (y + 0.5) / G, (z + 0.5) / G, 0.0))
    pos = np.array(pos, np.float32)
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    cfg = sd.default_config(particleCount=n, gridSize=G, softeningCells=soft, gravity=gravity, sortInterval=0)
    s = sd.Sim(cfg, sd.default_modes(exactGrid=1))
    s.set_particles(pos, vel)
    s.compute_forces()
    rho = s.get_grid(sd.GRID_DENSITY)
    pot = s.get_grid(sd.GRID_POTENTIAL)
    acc = s.get_grid(sd.GRID_ACCEL) if G <= 128 else None
    s.close()
    # the deposit puts each count on its node
    for x, y, z, m in src:
        assert rho[z, y, x] == pytest.approx(m, abs=0)
    assert float(rho.sum()) == pytest.approx(n, abs=0)
    ref = point_mass_potential(G, src, soft)
    scale = np.abs(ref).max()
    err = np.abs(pot.astype(np.float64) - ref).max() / scale
    _say("  pot rel max err", err)
    assert err < POT_TOL
    if acc is not None:
        potScale = np.float32(gravity) / np.float32(n)
        aref = oracle.grid_accel(ref, float(potScale), False)
        d = np.abs(acc[..., :3].astype(np.float64) - aref[..., :3]).max()
        h = 0.5 * G * float(potScale)
        _say("  acc max err", d, "of", np.abs(aref[..., :3]).max())
        # a central difference of two values within tol * scale
        assert d < 2 * POT_TOL * scale * h


@pytest.mark.parametrize("G", [128, 256, 512])
def test_periodic_plane_wave_closed_form(gpu, G):
    _say("periodic plane wave", G)
    b, c = 3, 7
    rng = np.random.default_rng(G)
    us = rng.choice(G, 10, replace=False)
    counts = {int(u): int(m) for u, m in zip(us, rng.integers(1, 4, 10))}
    pos = plane_particles(G, b, c, counts)
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    soft, gravity = 1.5, 0.9
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC, softeningCells=soft,
                            gravity=gravity, sortInterval=0)
    s = sd.Sim(cfg, sd.default_modes(exactGrid=1))
    s.set_particles(pos, vel)
    s.compute_forces()
    pot = s.get_grid(sd.GRID_POTENTIAL)
    acc = s.get_grid(sd.GRID_ACCEL) if G <= 256 else None
    s.close()
    g = plane_potential(G, b, c, counts, soft)
    scale = np.abs(g).max()
    x = np.arange(G)
    y = np.arange(G)
    err = 0.0
    aerr = 0.0
    h = float(np.float32(0.5) * np.float32(G) * (np.float32(gravity) / np.float32(n)))
    amax = 0.0
    for zz in range(G):
        u = (x[None, :] + b * y[:, None] + c * zz) % G
        err = max(err, float(np.abs(pot[zz].astype(np.float64) - g[u]).max()))
        if acc is not None:
            for axis, d in enumerate((1, b, c)):
                want = -(g[(u + d) % G] - g[(u - d) % G]) * h
                amax = max(amax, float(np.abs(want).max()))
                aerr = max(aerr, float(np.abs(acc[zz, :, :, axis].astype(np.float64) - want).max()))
    _say("  pot rel max err", err / scale, "acc max err", aerr, "of", amax)
    assert err / scale < POT_TOL
    if acc is not None:
        # a central difference of two values within tol * scale
        assert aerr < 2 * POT_TOL * scale * h
