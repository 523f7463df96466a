"""Initial conditions of the benchmark scenes, mapped into Stardust box units.

The reference's own scene generators (kPlace / Filament, src/sim/Sim.cu:102-481)
are not part of the PM path; the benchmark scenes are defined here:

* exponential disk: surface density ~ exp(-R/Rd) truncated at Rmax,
  vertical sech^2(z/h), circular speed sqrt(M(<R)/R) from the enclosed disk
  mass plus Gaussian dispersion (fraction of v_c per component), centre of
  mass at rest;
* two-disk encounter: the second disk inclined about x, offset and boosted;
* Zel'dovich lattice: displacements from a Gaussian random field with
  P(k) ~ k^n cut at the lattice Nyquist, scaled to an rms displacement.

Box mapping (the reference's units, Sim.cu:3107-3109, 563-611):
x_box = 0.5 + (x - centre)/L, v_box = v/L, time unchanged. Isolated meshes use
``gravity = G_mesh * G_N * M / L^3``; periodic meshes
``gravity = 4 pi G_mesh^3 * G_N * M / L^3``.
"""
from __future__ import annotations

import math

import numpy as np


def _disk_cdf(R, Rd):
    x = R / Rd
    return 1.0 - (1.0 + x) * np.exp(-x)


def exp_disk(n, seed=1, M=1.0, Rd=1.0, Rmax=5.0, h=0.1, disp=0.1, G_N=1.0):
    """Returns (pos[n,3], vel[n,3]) in physical units, COM at rest."""
    rng = np.random.default_rng(seed)
    grid = np.linspace(0.0, Rmax, 200001)
    cdf = _disk_cdf(grid, Rd)
    cmax = cdf[-1]
    u = rng.random(n) * cmax
    R = np.interp(u, cdf, grid)
    phi = rng.random(n) * 2.0 * math.pi
    uz = rng.random(n)
    uz = np.clip(uz, 1e-12, 1 - 1e-12)
    z = h * np.arctanh(2.0 * uz - 1.0)
    pos = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    Menc = M * _disk_cdf(R, Rd) / cmax
    vc = np.sqrt(G_N * Menc / np.maximum(R, 1e-12))
    vel = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(n)], axis=1)
    vel += rng.standard_normal((n, 3)) * (disp * vc)[:, None]
    vel -= vel.mean(axis=0)
    return pos, vel


def rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def two_disks(n_each, seed=1, incline_deg=45.0, offset=(8.0, 2.0, 0.0),
              rel_vel=(-0.5, 0.0, 0.0), **disk_kw):
    """Two disks as ``exp_disk``; the second inclined about x, offset and
    boosted. Returns (pos, vel, M_total) in the centre-of-mass frame."""
    p1, v1 = exp_disk(n_each, seed=seed, **disk_kw)
    p2, v2 = exp_disk(n_each, seed=seed + 1, **disk_kw)
    Rm = rot_x(math.radians(incline_deg))
    p2 = p2 @ Rm.T + np.asarray(offset)
    v2 = v2 @ Rm.T + np.asarray(rel_vel)
    pos = np.concatenate([p1, p2])
    vel = np.concatenate([v1, v2])
    vel -= vel.mean(axis=0)
    M = 2.0 * disk_kw.get("M", 1.0)
    return pos, vel, M


def to_box(pos, vel, centre, L):
    """float4 particle arrays in box units (w = 0)."""
    n = pos.shape[0]
    p4 = np.zeros((n, 4), np.float32)
    v4 = np.zeros((n, 4), np.float32)
    p4[:, :3] = 0.5 + (pos - np.asarray(centre)) / L
    v4[:, :3] = vel / L
    return p4, v4


def gravity_isolated(G_mesh, M, L, G_N=1.0):
    return float(G_mesh * G_N * M / L ** 3)


def gravity_periodic(G_mesh, M, L, G_N=1.0):
    return float(4.0 * math.pi * G_mesh ** 3 * G_N * M / L ** 3)


def softening_cells(G_mesh):
    """The application's softening policy, softeningCells = 3 G / 128
    (src/app/App.cpp:483): a fixed physical length at every resolution."""
    return 3.0 * G_mesh / 128.0


# ---- benchmark scenes ---------------------------------------------------------

def config0(n=262144, seed=1):
    """Exponential disk, G=64 / S=128 isolated, dt 1e-3, 10 steps."""
    pos, vel = exp_disk(n, seed=seed)
    L = 12.0
    p4, v4 = to_box(pos, vel, (0.0, 0.0, 0.0), L)
    return dict(pos=p4, vel=v4, G=64, L=L, M=1.0, dt=1e-3, steps=10,
                gravity=gravity_isolated(64, 1.0, L), soft=softening_cells(64))


def config1(n_each=2097152, seed=1, G=128):
    """Two-disk encounter, G=128 / S=256 isolated, 100 steps, re-sort every 10."""
    pos, vel, M = two_disks(n_each, seed=seed)
    L = 20.0
    centre = (4.0, 0.3, 0.0)
    p4, v4 = to_box(pos, vel, centre, L)
    return dict(pos=p4, vel=v4, G=G, L=L, M=M, dt=1e-3, steps=100, sort_interval=10,
                gravity=gravity_isolated(G, M, L), soft=softening_cells(G))


def zeldovich(n_side=256, seed=1, rms_cells=0.3, G_mesh=512, index=-2.0):
    """Lattice of n_side^3 particles displaced by a Gaussian random field
    (P(k) ~ k^index up to the lattice Nyquist) with rms displacement
    ``rms_cells`` mesh cells; periodic box of side 1, velocities zero."""
    rng = np.random.default_rng(seed)
    ns = n_side
    k = np.fft.fftfreq(ns) * ns
    kz = k[:, None, None]
    ky = k[None, :, None]
    kx = np.fft.rfftfreq(ns)[None, None, :] * ns
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    amp = np.sqrt(k2 ** (index / 2.0))
    amp[0, 0, 0] = 0.0
    amp[np.sqrt(k2) > ns / 2] = 0.0
    phase = rng.standard_normal((ns, ns, ns // 2 + 1)) + 1j * rng.standard_normal((ns, ns, ns // 2 + 1))
    delta_k = (amp * phase).astype(np.complex64)
    disp = []
    for kk in (kx, ky, kz):
        psi_k = (1j * kk / k2) * delta_k
        disp.append(np.fft.irfftn(psi_k, s=(ns, ns, ns), axes=(0, 1, 2)).astype(np.float32))
    d = np.stack([disp[0], disp[1], disp[2]], axis=-1)
    rms = float(np.sqrt((d.astype(np.float64) ** 2).sum(axis=-1).mean()))
    d *= (rms_cells / G_mesh) / max(rms, 1e-30)
    g = (np.arange(ns, dtype=np.float32) + 0.5) / ns
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    n = ns ** 3
    p4 = np.zeros((n, 4), np.float32)
    p4[:, 0] = (xx + d[..., 0]).reshape(-1)
    p4[:, 1] = (yy + d[..., 1]).reshape(-1)
    p4[:, 2] = (zz + d[..., 2]).reshape(-1)
    p4[:, :3] -= np.floor(p4[:, :3])
    v4 = np.zeros((n, 4), np.float32)
    return p4, v4


def config3(n_scenes=64, seed=3, n_range=(2 ** 15, 2 ** 20), meshes=(64, 96, 128)):
    """Batched ensemble: ``n_scenes`` independent disks as config 0, particle
    counts N_i = round(exp(U[ln lo, ln hi])), meshes G_i drawn uniformly from
    ``meshes`` (S = 2 G), disk seed = scene index. Returns a list of
    (n, G, seed) triples; ``disk_scene`` builds each."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    out = []
    for i in range(n_scenes):
        n = int(round(math.exp(rng.uniform(lo, hi))))
        G = int(meshes[int(rng.integers(len(meshes)))])
        out.append((n, G, i))
    return out


def disk_scene(n, G, seed, L=12.0):
    """One config-0 disk of n particles on a G mesh (box side L, dt 1e-3)."""
    pos, vel = exp_disk(n, seed=seed)
    p4, v4 = to_box(pos, vel, (0.0, 0.0, 0.0), L)
    return dict(pos=p4, vel=v4, G=G, L=L, M=1.0, dt=1e-3,
                gravity=gravity_isolated(G, 1.0, L), soft=softening_cells(G))
