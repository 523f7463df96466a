"""Closed forms of the mesh solves, shared by the oracle tests (where they
are checked against the oracle's FP64 solves at small sizes) and the GPU
tests (where they stand in for the oracle at the large sizes).

* Isolated: point masses at cell centres put their whole count on one node
  each, and the potential on the G^3 corner of the zero-padded convolution is
  sum_k -m_k / (r_k + soft) at every node (kGreen, Sim.cu:563-573; the
  padding leaves no periodic images).
* Periodic: a density that depends only on u = (x + b y + c z) mod G is a sum
  of plane waves along (1, b, c); its potential is g(u), the 1-D inverse DFT
  of the density's 1-D spectrum times the reference multiplier
  -1/(4 pi^2 k^2) exp(-k^2 s^2 / 2), s = 2 pi soft / G (kPoissonPeriodic,
  Sim.cu:528-557), with k the folded wavevector j (1, b, c) mod G.
"""
import numpy as np


def point_mass_potential(G, src, soft):
    """Isolated potential of point masses src = [(x, y, z, count)] on the G^3
    nodes [z][y][x], with soft = max(softeningCells, 0.5) as kGreen."""
    soft = max(float(soft), 0.5)
    ax = np.arange(G, dtype=np.float64)
    ref = np.zeros((G, G, G))
    for x, y, z, m in src:
        r = np.sqrt((ax[None, None, :] - x) ** 2 + (ax[None, :, None] - y) ** 2 + (ax[:, None, None] - z) ** 2)
        ref -= m / (r + soft)
    return ref


def point_masses(G):
    """(x, y, z, count) of a few sources: mesh corners, an edge, the centre."""
    return [(0, 0, 0, 1), (G - 1, G - 1, G - 1, 2), (G - 1, 0, G // 2, 1), (G // 3, 2 * G // 3, 5, 3),
            (G // 2, G // 2, G // 2, 1), (1, G - 2, G - 1, 2)]


def plane_particles(G, b, c, counts):
    """Particles at cell centres with count counts[u] in every cell whose
    u = (x + b y + c z) mod G; returns float32 [n, 4]."""
    yz = np.arange(G * G)
    y, z = yz % G, yz // G
    out = []
    for u, m in counts.items():
        x = (u - b * y - c * z) % G
        p = np.stack([(x + 0.5) / G, (y + 0.5) / G, (z + 0.5) / G, np.zeros(G * G)], axis=1).astype(np.float32)
        out.extend([p] * m)
    return np.concatenate(out, axis=0)


def plane_potential(G, b, c, counts, soft):
    """g(u) of the periodic solve for the plane-wave density (FP64)."""
    f = np.zeros(G)
    for u, m in counts.items():
        f[u] = m
    F = np.fft.fft(f)
    j = np.arange(G)
    k2 = np.zeros(G)
    for comp in (1, b, c):
        k = (j * comp) % G
        k = np.where(k <= G // 2, k, k - G)
        k2 += k.astype(np.float64) ** 2
    s = 2.0 * np.pi * soft / G
    mult = np.zeros(G)
    nz = k2 > 0
    mult[nz] = -1.0 / (4.0 * np.pi ** 2 * k2[nz]) * np.exp(-0.5 * k2[nz] * s * s)
    # the 3-D spectrum of the density is G^2 F_j on the wavevector j (1, b, c);
    # the normalised inverse 3-D transform of G^2 F_j mult_j e^{..} is g(u)
    return np.fft.ifft(F * mult).real
