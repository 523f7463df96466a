"""Batched scenes (one launch per kernel for all scenes, a run recorded into
one HIP graph) against the same scenes stepped alone: bit-identical positions,
velocities, times and snapshots, for uneven particle counts, 2^k and 3*2^k
meshes, both boundaries and both dt policies."""
import numpy as np
import pytest

import stardust_amd as sd
from stardust_amd import ic

SORT = 5


def _scenes():
    """(config, pos, vel, fixed dt) of a small heterogeneous batch."""
    out = []
    for n, G, seed, periodic, dt in [(3000, 16, 11, 0, 1e-3), (20000, 32, 12, 0, 1e-3), (50000, 32, 13, 0, 0.0),
                                     (7000, 16, 14, 0, 2e-3), (30000, 96, 15, 0, 1e-3), (1, 16, 16, 0, 1e-3),
                                     (12000, 32, 17, 1, 1e-3)]:
        d = ic.disk_scene(n, G, seed)
        grav = ic.gravity_periodic(G, 1.0, d["L"]) if periodic else d["gravity"]
        cfg = sd.default_config(particleCount=n, gridSize=G, gravity=grav, softeningCells=d["soft"],
                                sortInterval=SORT, boundary=periodic)
        out.append((cfg, d["pos"], d["vel"], dt))
    # inactive slots (x < 0) among live ones: a scene stepped alone reorders
    # inside its integrate launch, the batch with a separate gather
    d = ic.disk_scene(6000, 32, 18)
    pos, vel = d["pos"].copy(), d["vel"].copy()
    dead = np.random.default_rng(4).random(6000) < 0.2
    pos[dead] = (-1.0, -1.0, -1.0, 0.0)
    vel[dead] = 0.0
    cfg = sd.default_config(particleCount=6000, gridSize=32, gravity=d["gravity"], softeningCells=d["soft"],
                            sortInterval=SORT)
    out.append((cfg, pos, vel, 1e-3))
    return out


def _batch(scenes):
    b = sd.Batch([c for c, _, _, _ in scenes])
    for i, (_, p, v, dt) in enumerate(scenes):
        b.scene(i).set_particles(p, v)
        b.scene(i).set_fixed_dt(dt)
    return b


def _alone(cfg, p, v, dt):
    s = sd.Sim(cfg, sd.default_modes(exactGrid=1))
    s.set_particles(p, v)
    s.set_fixed_dt(dt)
    return s


@pytest.mark.gpu
def test_batch_matches_scenes_stepped_alone(gpu):
    scenes = _scenes()
    b = _batch(scenes)
    steps, every = 12, 4
    b.run(steps, every)
    b.synchronize()
    assert b.graph_nodes() > 0 and b.snapshot_count() == steps // every
    for i, (cfg, p, v, dt) in enumerate(scenes):
        s = _alone(cfg, p, v, dt)
        for k in range(steps // every):
            s.run(every)
            pa, _ = s.get_particles()
            np.testing.assert_array_equal(b.snapshot(k, i), pa, err_msg=f"scene {i} snapshot {k}")
        pa, va = s.get_particles()
        pb, vb = b.scene(i).get_particles()
        np.testing.assert_array_equal(pb, pa, err_msg=f"scene {i} positions")
        np.testing.assert_array_equal(vb, va, err_msg=f"scene {i} velocities")
        assert b.scene(i).sim_time() == s.sim_time()
        live = int((p[:, 0] >= 0).sum())
        assert b.scene(i).measure_total_grid_mass() == pytest.approx(live, abs=1e-6)
        s.close()
    b.close()


@pytest.mark.gpu
def test_batch_replay_continues_like_single_scenes(gpu):
    """A second run of a whole number of sort intervals replays the recorded
    graph; the scenes continue exactly as scenes stepped alone for 2 x 10."""
    scenes = _scenes()[:4]
    b = _batch(scenes)
    b.run(10, 0)
    nodes = b.graph_nodes()
    b.run(10, 0)
    b.synchronize()
    assert b.graph_nodes() == nodes
    for i, (cfg, p, v, dt) in enumerate(scenes):
        s = _alone(cfg, p, v, dt)
        s.run(20)
        pa, va = s.get_particles()
        pb, vb = b.scene(i).get_particles()
        np.testing.assert_array_equal(pb, pa, err_msg=f"scene {i}")
        np.testing.assert_array_equal(vb, va, err_msg=f"scene {i}")
        assert b.scene(i).step_count() == 20
        s.close()
    b.close()


@pytest.mark.gpu
def test_batch_scene_calls_and_refusals(gpu):
    scenes = _scenes()[:3]
    b = _batch(scenes)
    sc = b.scene(1)
    # a scene keeps its size; structural reconfigure is refused
    cfg = scenes[1][0]
    bigger = sd.default_config(particleCount=cfg.particleCount + 1, gridSize=cfg.gridSize, gravity=cfg.gravity,
                               softeningCells=cfg.softeningCells, sortInterval=SORT)
    with pytest.raises(sd.StardustError):
        sc.reconfigure(bigger)
    assert not sd.lib().sd_failed()
    # a scene-level sort keeps the batch consistent (copy back, same addresses)
    sc.sort_now()
    b.run(3, 0)
    b.synchronize()
    for i in range(3):
        assert b.scene(i).measure_total_grid_mass() == pytest.approx(scenes[i][0].particleCount, abs=1e-6)
    b.close()


@pytest.mark.gpu
def test_batch_config3_meshes_vs_oracle(gpu, oracle):
    """Batched disks on config 3's meshes (G = 64, 96 and 128, S = 2G) against
    the oracle: each scene's leapfrog from v0 equals the oracle's kick-drift
    (Sim::step, Morton keys) from v0 - a(x0) dt / 2, over a run that crosses a
    sort (steps 0 and 5)."""
    steps, dt = 7, 1e-3
    specs = [(40000, 64, 21), (30000, 96, 22), (50000, 128, 23), (9000, 64, 24)]
    scenes = []
    for n, G, seed in specs:
        d = ic.disk_scene(n, G, seed)
        pos = d["pos"].copy()
        pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)
        cfg = sd.default_config(particleCount=n, gridSize=G, gravity=d["gravity"], softeningCells=d["soft"],
                                sortInterval=SORT)
        scenes.append((cfg, pos, d["vel"], dt))
    b = _batch(scenes)
    b.run(steps, 0)
    b.synchronize()
    for i, (cfg, pos, vel, _) in enumerate(scenes):
        n, G = cfg.particleCount, cfg.gridSize
        print("[case] batch scene vs oracle", i, n, G, flush=True)
        gp, _ = b.scene(i).get_particles()
        gp = gp[np.argsort(gp[:, 3])]
        prm = oracle.make_params(G, gravity=cfg.gravity, softeningCells=cfg.softeningCells, sortInterval=SORT,
                                 keyMode=oracle.KEY_MORTON, fixedDt=dt, allocN=n)
        a0 = oracle.particle_accel(prm, pos)
        v0 = vel.copy()
        v0[:, :3] -= 0.5 * dt * a0[:, :3]
        st = oracle.Stepper(prm, pos, v0)
        st.step(steps)
        op = st.pos[np.argsort(st.pos[:, 3])]
        assert np.array_equal(gp[:, 3], op[:, 3])
        dp = np.abs(gp[:, :3] - op[:, :3]).max()
        print("[case]   max |dx|", dp, flush=True)
        assert dp < 1e-6  # measured 6e-8 (one ulp at 0.5-1)
        assert b.scene(i).sim_time() == pytest.approx(st.state.simTime, rel=1e-6)
    b.close()


@pytest.mark.gpu
def test_batch_mode_change_rerecords(gpu):
    """The recorded graph bakes in each scene's modes (integrator, momentum
    removal, key order); changing them between runs must not replay the old
    recording. Against scenes stepped alone with the same change."""
    scenes = _scenes()[:3]
    b = _batch(scenes)
    b.run(SORT, 0)
    kd = sd.default_modes(exactGrid=1, integrator=sd.INTEGRATOR_KD, momentumZero=0)
    b.scene(1).set_modes(kd)
    b.run(SORT, 0)
    b.synchronize()
    for i, (cfg, p, v, dt) in enumerate(scenes):
        s = _alone(cfg, p, v, dt)
        s.run(SORT)
        if i == 1:
            s.set_modes(kd)
        s.run(SORT)
        pa, va = s.get_particles()
        pb, vb = b.scene(i).get_particles()
        np.testing.assert_array_equal(pb, pa, err_msg=f"scene {i}")
        np.testing.assert_array_equal(vb, va, err_msg=f"scene {i}")
        s.close()
    b.close()
