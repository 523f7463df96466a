"""The benchmark configurations as parity cases against the oracle.

* config 0 at its full size (262,144-particle disk, G=64 / S=128 isolated,
  dt 1e-3, 10 steps): the reference-order pipeline against the oracle's
  Sim::step, and the product pipeline (fused, leapfrog, Morton keys) against
  the oracle's kick-drift started half a kick back (the exact relation of the
  two integrators);
* config 2 at reduced size (Zel'dovich lattice 32^3 on a periodic G=64 mesh,
  contact radius 0.2 cells, re-sort every step, CFL dt).

Tolerances: positions 5e-6 box units, velocities 2e-4 of the largest speed
(FP32 FFT on the GPU against the oracle's FP64 FFT).
"""
import sys

import numpy as np
import pytest

import stardust_amd as sd
from stardust_amd import ic

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _config0():
    c = ic.config0()
    pos = c["pos"].copy()
    pos[:, 3] = np.arange(1, pos.shape[0] + 1, dtype=np.float32)  # identity in w (unused by the step)
    return c, pos, c["vel"]


def test_config0_reference_order_vs_oracle(gpu, oracle):
    c, pos, vel = _config0()
    n, G, steps = pos.shape[0], c["G"], c["steps"]
    _say("config0 reference order", n, G, steps)
    cfg = sd.default_config(particleCount=n, gridSize=G, gravity=c["gravity"], softeningCells=c["soft"],
                            sortInterval=40)
    s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD,
                                     keyMode=sd.KEY_ROWMAJOR))
    s.set_fixed_dt(c["dt"])
    s.set_particles(pos, c["vel"])
    s.run(steps)
    gp, gv = s.get_particles()
    t = s.sim_time()
    s.close()
    prm = oracle.make_params(G, gravity=c["gravity"], softeningCells=c["soft"], sortInterval=40,
                             keyMode=oracle.KEY_ROWMAJOR, fixedDt=c["dt"], allocN=n)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    assert np.array_equal(gp[:, 3], st.pos[:, 3])  # same permutation
    dp = np.abs(gp[:, :3] - st.pos[:, :3]).max()
    dv = np.abs(gv[:, :3] - st.vel[:, :3]).max() / np.abs(st.vel[:, :3]).max()
    _say("  max |dx|", dp, "max rel dv", dv)
    assert dp < 5e-6
    assert dv < 2e-4
    assert abs(t - st.state.simTime) < 1e-9 + 1e-6 * st.state.simTime


@pytest.mark.parametrize("interval,merge", [(40, 0), (3, 1)])
def test_config0_product_path_vs_oracle(gpu, oracle, interval, merge):
    """merge = 1: every re-sort after the first starts from the last sort's
    keys (the changed keys sorted alone, then merged)."""
    c, pos, vel = _config0()
    n, G, steps, dt = pos.shape[0], c["G"], c["steps"], c["dt"]
    _say("config0 product path", n, G, steps, "sort every", interval, "merge", merge)
    cfg = sd.default_config(particleCount=n, gridSize=G, gravity=c["gravity"], softeningCells=c["soft"],
                            sortInterval=interval)
    s = sd.Sim(cfg, sd.default_modes(sortMerge=merge))  # fused, leapfrog, Morton keys
    s.set_fixed_dt(dt)
    s.set_particles(pos, vel)
    s.run(steps)
    gp, gv = s.get_particles()
    s.close()
    # leapfrog from v0 = kick-drift from v0 - a(x0) dt / 2 (positions); a(x0) from the oracle
    prm = oracle.make_params(G, gravity=c["gravity"], softeningCells=c["soft"], sortInterval=interval,
                             keyMode=oracle.KEY_MORTON, fixedDt=dt, allocN=n)
    a0 = oracle.particle_accel(prm, pos)
    v0 = vel.copy()
    v0[:, :3] -= 0.5 * dt * a0[:, :3]
    st = oracle.Stepper(prm, pos, v0)
    st.step(steps)
    og = np.argsort(gp[:, 3])
    oo = np.argsort(st.pos[:, 3])
    dp = np.abs(gp[og, :3] - st.pos[oo, :3]).max()
    _say("  max |dx|", dp)
    assert np.array_equal(gp[og, 3], st.pos[oo, 3])
    assert dp < 5e-6


@pytest.mark.parametrize("key", ["rowmajor", "morton"])
def test_config2_small_periodic_contact_vs_oracle(gpu, oracle, key):
    """Row-major keys read the per-cell table; Morton keys find each cell's
    particles inside its 4^3-cell block's run of the sorted keys."""
    pos, vel = ic.zeldovich(32, seed=1, rms_cells=0.3, G_mesh=64)
    n, G, steps = pos.shape[0], 64, 4
    grav = ic.gravity_periodic(G, 1.0, 1.0) * 1e-9  # weak gravity: contact and mesh both matter
    skey, okey = (sd.KEY_ROWMAJOR, oracle.KEY_ROWMAJOR) if key == "rowmajor" else (sd.KEY_MORTON, oracle.KEY_MORTON)
    _say("config2 small", key, n, G, steps)
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC, gravity=grav,
                            softeningCells=3.0, sortInterval=1, contactEnabled=1, contactRadiusCells=0.2,
                            contactStiffness=1e3)
    vel = vel.copy()
    vel[:, :3] = np.random.default_rng(2).normal(0, 0.002, (n, 3)).astype(np.float32)
    s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD,
                                     keyMode=skey))
    s.set_particles(pos, vel)
    s.run(steps)
    gp, gv = s.get_particles()
    s.close()
    prm = oracle.make_params(G, periodic=True, gravity=grav, softeningCells=3.0, sortInterval=1,
                             keyMode=okey, contact=1, contactRadiusCells=0.2,
                             contactStiffness=1e3, allocN=n)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    dp = np.abs(gp[:, :3] - st.pos[:, :3]).max()
    dv = np.abs(gv[:, :3] - st.vel[:, :3]).max() / np.abs(st.vel[:, :3]).max()
    _say("  max |dx|", dp, "max rel dv", dv)
    assert dp < 5e-6
    assert dv < 2e-4


def _product_vs_oracle(oracle, pos, vel, G, gravity, soft, dt, steps, sort_interval, periodic=False):
    """Leapfrog from v0 on the GPU against the oracle's kick-drift from
    v0 - a(x0) dt / 2 (the exact relation of the two integrators); returns the
    GPU and oracle particles ordered by the identity in w."""
    n = pos.shape[0]
    prm = oracle.make_params(G, periodic=periodic, gravity=gravity, softeningCells=soft,
                             sortInterval=sort_interval, keyMode=oracle.KEY_MORTON, fixedDt=dt, allocN=n)
    a0 = oracle.particle_accel(prm, pos)
    v0 = vel.copy()
    v0[:, :3] -= 0.5 * dt * a0[:, :3]
    st = oracle.Stepper(prm, pos, v0)
    st.step(steps)
    og = np.argsort(st.pos[:, 3])
    return st.pos[og], st


def test_config1_full_size_product_path_vs_oracle(gpu, oracle):
    """The headline workload at its full size: two-disk encounter, 4,194,304
    particles, G=128 / S=256 isolated, dt 1e-3, Morton re-sort every 10 steps
    (a sort on the first step). Product path (fused kernels, leapfrog, Morton
    keys) for 4 steps against the oracle's Sim::step."""
    c = ic.config1()
    pos = c["pos"].copy()
    n, G, dt, steps = pos.shape[0], c["G"], c["dt"], 4
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)
    _say("config1 full size", n, G, steps)
    cfg = sd.default_config(particleCount=n, gridSize=G, gravity=c["gravity"], softeningCells=c["soft"],
                            sortInterval=c["sort_interval"])
    s = sd.Sim(cfg, sd.default_modes())
    s.set_fixed_dt(dt)
    s.set_particles(pos, c["vel"])
    s.run(steps)
    gp, _ = s.get_particles()
    s.close()
    gp = gp[np.argsort(gp[:, 3])]
    op, _ = _product_vs_oracle(oracle, pos, c["vel"], G, c["gravity"], c["soft"], dt, steps, c["sort_interval"])
    assert np.array_equal(gp[:, 3], op[:, 3])
    dp = np.abs(gp[:, :3] - op[:, :3]).max()
    _say("  max |dx|", dp)
    assert dp < 1e-6  # measured 6e-8 (one ulp at 0.5-1)


def test_config2_full_mesh_periodic_contact_vs_oracle(gpu, oracle):
    """Config 2's mesh at its full size: periodic G=512, contact radius 0.2
    cells, re-sort every step, CFL dt, on a reduced lattice (128^3 particles
    with Zel'dovich displacements, rms 0.3 cells) plus a dense clump of 20,000
    particles within 1.5 cells that reaches the 96-candidate cap. Reference
    kernel order and the fused pipeline (both kick-drift, Morton keys) against
    the oracle's Sim::step for 2 steps."""
    G, steps = 512, 2
    pos, _ = ic.zeldovich(128, seed=3, rms_cells=0.3, G_mesh=G)
    n = pos.shape[0]
    rng = np.random.default_rng(5)
    m = 20000
    clump = np.array([0.31, 0.62, 0.47]) + rng.normal(0, 0.7 / G, (m, 3))
    pos[:m, :3] = np.mod(clump, 1.0).astype(np.float32)
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.002, (n, 3)).astype(np.float32)
    grav = ic.gravity_periodic(G, 1.0, 1.0)
    _say("config2 full mesh", n, G, steps)
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC, gravity=grav,
                            softeningCells=3.0, sortInterval=1, contactEnabled=1, contactRadiusCells=0.2,
                            contactStiffness=1e3)
    runs = []
    for pipe in (sd.PIPELINE_REFERENCE, sd.PIPELINE_FUSED):
        s = sd.Sim(cfg, sd.default_modes(pipeline=pipe, integrator=sd.INTEGRATOR_KD, keyMode=sd.KEY_MORTON))
        s.set_particles(pos, vel)
        s.run(steps)
        p, v = s.get_particles()
        t = s.sim_time()
        s.close()
        o = np.argsort(p[:, 3])
        runs.append((p[o], v[o], t))
    prm = oracle.make_params(G, periodic=True, gravity=grav, softeningCells=3.0, sortInterval=1,
                             keyMode=oracle.KEY_MORTON, contact=1, contactRadiusCells=0.2,
                             contactStiffness=1e3, allocN=n)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    o = np.argsort(st.pos[:, 3])
    op, ov = st.pos[o], st.vel[o]
    vs = np.abs(ov[:, :3]).max()
    assert vs > 0.01  # the clump's contacts act
    for (gp, gv, t), name in zip(runs, ("reference order", "fused")):
        assert np.array_equal(gp[:, 3], op[:, 3])
        dp = np.abs(gp[:, :3] - op[:, :3]).max()
        dv = np.abs(gv[:, :3] - ov[:, :3]).max() / vs
        _say(" ", name, "max |dx|", dp, "max rel dv", dv, "t", t, st.state.simTime)
        assert dp < 1e-6  # measured 6e-8 (one ulp at 0.5-1)
        assert dv < 2e-4
        assert abs(t - st.state.simTime) < 1e-6 * st.state.simTime
