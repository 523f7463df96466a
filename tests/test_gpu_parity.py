"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: bit-exact for keys, sort permutation and the fixed-point density; FP32
tolerances (stated per test) for the FFT potential, accelerations and
trajectories, where the oracle solves the mesh in FP64.
"""
import sys

import numpy as np
import pytest

import stardust_amd as sd

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def random_particles(n, seed, inactive=0.05, lo=0.05, hi=0.95, clustered=True):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    if clustered:
        # half uniform, half in a few tight clumps (exercises contention)
        m = n // 2
        pos[:m, :3] = rng.uniform(lo, hi, (m, 3))
        centres = rng.uniform(0.3, 0.7, (4, 3))
        pick = rng.integers(0, 4, n - m)
        pos[m:, :3] = centres[pick] + rng.normal(0, 0.02, (n - m, 3))
    else:
        pos[:, :3] = rng.uniform(lo, hi, (n, 3))
    pos[:, :3] = np.clip(pos[:, :3], 0.003, 0.997)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.05, (n, 3))
    dead = rng.random(n) < inactive
    pos[dead] = (-1, -1, -1, 0)
    vel[dead] = 0
    return pos, vel


def make_sim(n, G, periodic=False, soft=1.5, gravity=1.0, sort_interval=0, **modes):
    cfg = sd.default_config(particleCount=n, gridSize=G,
                            boundary=sd.BOUNDARY_PERIODIC if periodic else sd.BOUNDARY_ISOLATED,
                            softeningCells=soft, gravity=gravity, sortInterval=sort_interval)
    m = sd.default_modes(**modes)
    return sd.Sim(cfg, m)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("key_mode", [sd.KEY_ROWMAJOR, sd.KEY_MORTON])
@pytest.mark.parametrize("n,G", [(1, 16), (777, 16), (20000, 32), (300000, 128)])
def test_sort_bit_exact(gpu, oracle, key_mode, n, G):
    _say("sort", key_mode, n, G)
    pos, vel = random_particles(n, seed=n + G)
    s = make_sim(n, G, keyMode=key_mode)
    s.set_particles(pos, vel)
    keys, order = s.sort_now()
    # the step sort of Morton mode orders by the CIC base node
    ok, oi = oracle.cell_keys(pos, G, oracle.KEY_MORTON_NODE if key_mode else oracle.KEY_ROWMAJOR)
    sk, so = oracle.stable_sort_pairs(ok, oi)
    assert np.array_equal(keys, sk)
    assert np.array_equal(order, so)
    p2, v2 = s.get_particles()
    assert np.array_equal(p2, pos[so])
    s.close()


@pytest.mark.parametrize("key_mode", [sd.KEY_ROWMAJOR, sd.KEY_MORTON])
@pytest.mark.parametrize("n,G,frac", [(5000, 16, 0.0), (20000, 32, 0.03), (300000, 128, 0.1),
                                      (300000, 128, 1.0), (1000003, 64, 0.3)])
def test_resort_of_nearly_sorted_keys(gpu, oracle, key_mode, n, G, frac):
    """A second sort starts from the first one's ascending keys: the keys that
    changed are sorted alone and merged into the unchanged ones. The result
    must be the stable sort of all keys, for any fraction of changed keys
    (frac of the particles jump anywhere, the rest move by up to 0.2 cells),
    and also when the previous keys belong to other particles entirely."""
    _say("resort", key_mode, n, G, frac)
    pos, vel = random_particles(n, seed=n + G + 1)
    s = make_sim(n, G, keyMode=key_mode, sortMerge=1)
    s.set_particles(pos, vel)
    s.sort_now()
    p1, v1 = s.get_particles()
    rng = np.random.default_rng(n + 7)
    live = p1[:, 0] >= 0
    p2 = p1.copy()
    p2[live, :3] += rng.uniform(-0.2, 0.2, (int(live.sum()), 3)).astype(np.float32) / G
    jump = live & (rng.random(n) < frac)
    p2[jump, :3] = rng.uniform(0.01, 0.99, (int(jump.sum()), 3)).astype(np.float32)
    p2[:, :3] = np.where(live[:, None], np.clip(p2[:, :3], 0.003, 0.997), p2[:, :3])
    mode = oracle.KEY_MORTON_NODE if key_mode else oracle.KEY_ROWMAJOR
    for label, p in (("moved", p2), ("unrelated", random_particles(n, seed=n + G + 2)[0])):
        s.set_particles(p, v1)
        keys, order = s.sort_now()
        ok, oi = oracle.cell_keys(p, G, mode)
        sk, so = oracle.stable_sort_pairs(ok, oi)
        _say("  ", label)
        assert np.array_equal(keys, sk), label
        assert np.array_equal(order, so), label
        q, _ = s.get_particles()
        assert np.array_equal(q, p[so]), label
    s.close()


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("n,G", [(5000, 16), (100000, 32), (400000, 64)])
def test_deposit_fixed_bit_exact(gpu, oracle, periodic, n, G):
    _say("deposit", periodic, n, G)
    pos, vel = random_particles(n, seed=3 * n + G, lo=-0.02 if periodic else 0.05)
    if periodic:
        pos[:, :3] = np.where(pos[:, 0:1] >= 0, np.mod(pos[:, :3], 1.0), pos[:, :3])
    s = make_sim(n, G, periodic=periodic)
    s.set_particles(pos, vel)
    rho = s.get_grid(sd.GRID_DENSITY_FIXED)
    ref = oracle.deposit_fixed(pos, G, periodic)
    assert np.array_equal(rho, ref)
    live = int((pos[:, 0] >= 0).sum())
    assert int(rho.sum()) == live * sd.FIX_ONE  # exact mass
    # against the reference's float deposit (order-dependent): its sequential
    # FP32 sums carry ~count * 2^-24 relative error, the fixed point 2^-24 per particle
    rf = oracle.deposit_float(pos, G, G, periodic)
    assert np.abs(rho / sd.FIX_ONE - rf).max() < 1e-5 * max(1.0, rf.max())
    s.close()


@pytest.mark.parametrize("periodic,G", [(False, 8), (False, 16), (False, 32), (False, 64), (False, 96),
                                        (False, 128), (True, 16), (True, 32), (True, 64), (True, 128)])
def test_potential_and_accel(gpu, oracle, periodic, G):
    _say("mesh", periodic, G)
    n = 20000
    pos, vel = random_particles(n, seed=G + 11)
    if periodic:
        pos[:, :3] = np.where(pos[:, 0:1] >= 0, np.mod(pos[:, :3], 1.0), pos[:, :3])
    soft = 1.5
    gravity = 0.8
    s = make_sim(n, G, periodic=periodic, soft=soft, gravity=gravity, exactGrid=1)
    s.set_particles(pos, vel)
    s.compute_forces()
    pot = s.get_grid(sd.GRID_POTENTIAL).astype(np.float64)
    acc = s.get_grid(sd.GRID_ACCEL)
    rho = oracle.deposit_fixed(pos, G, periodic).astype(np.float64) / oracle.FIX_ONE
    if periodic:
        pref = oracle.poisson_periodic(rho, soft)
    else:
        pref = oracle.poisson_iso(rho, soft)
    scale = np.abs(pref).max()
    err = np.abs(pot - pref).max() / scale
    _say("  pot rel max err", err)
    assert err < 2e-5  # FP32 transforms vs FP64
    potScale = np.float32(gravity) / np.float32(n)
    aref = oracle.grid_accel(pref, float(potScale), periodic)
    d = np.linalg.norm(acc[..., :3] - aref[..., :3], axis=-1)
    rms = np.sqrt((d ** 2).mean()) / np.sqrt((np.linalg.norm(aref[..., :3], axis=-1) ** 2).mean())
    _say("  acc rel rms err", rms)
    assert rms < 1e-4
    s.close()


def test_potential_point_mass(gpu):
    """Isolated Green's function: a point mass at a cell centre gives
    pot = -1/(r_cells + soft) exactly at every node (no periodic images)."""
    G = 32
    n = 1
    pos = np.array([[(10 + 0.5) / G, (12 + 0.5) / G, (9 + 0.5) / G, 0]], np.float32)
    vel = np.zeros((1, 4), np.float32)
    s = make_sim(n, G, soft=2.0)
    s.set_particles(pos, vel)
    pot = s.get_grid(sd.GRID_POTENTIAL).astype(np.float64)
    z, y, x = np.meshgrid(np.arange(G), np.arange(G), np.arange(G), indexing="ij")
    r = np.sqrt((x - 10.0) ** 2 + (y - 12.0) ** 2 + (z - 9.0) ** 2)
    ref = -1.0 / (r + 2.0)
    assert np.abs(pot - ref).max() < 2e-6
    s.close()


def _run_both(pos, vel, G, steps, periodic=False, soft=1.5, gravity=0.5, fixed_dt=0.0,
              sort_interval=3, pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD,
              key_mode=sd.KEY_ROWMAJOR, momentum=1, oracle=None):
    n = pos.shape[0]
    s = make_sim(n, G, periodic=periodic, soft=soft, gravity=gravity, sort_interval=sort_interval,
                 pipeline=pipeline, integrator=integrator, keyMode=key_mode, momentumZero=momentum)
    s.set_fixed_dt(fixed_dt)
    s.set_particles(pos, vel)
    s.run(steps)
    gp, gv = s.get_particles()
    t = s.sim_time()
    s.close()
    prm = oracle.make_params(G, periodic=periodic, gravity=gravity, softeningCells=soft,
                             sortInterval=sort_interval, keyMode=key_mode, fixedDt=fixed_dt,
                             momentumZero=momentum, allocN=n)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    return gp, gv, t, st


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("fixed_dt", [0.0, 2e-3])
def test_reference_step_vs_oracle(gpu, oracle, periodic, fixed_dt):
    _say("step", periodic, fixed_dt)
    n, G, steps = 6000, 32, 6
    pos, vel = random_particles(n, seed=5)
    if periodic:
        pos[:, :3] = np.where(pos[:, 0:1] >= 0, np.mod(pos[:, :3], 1.0), pos[:, :3])
    gp, gv, t, st = _run_both(pos, vel, G, steps, periodic=periodic, fixed_dt=fixed_dt,
                              oracle=oracle)
    # same permutation (row-major keys, same sort schedule)
    assert np.array_equal(gp[:, 0] < 0, st.pos[:, 0] < 0)
    dp = np.abs(gp[:, :3] - st.pos[:, :3]).max()
    vs = np.abs(st.vel[:, :3]).max()
    dv = np.abs(gv[:, :3] - st.vel[:, :3]).max() / vs
    _say("  max |dx|", dp, "max rel dv", dv)
    assert dp < 2e-6
    assert dv < 1e-4
    assert abs(t - st.state.simTime) < 1e-6 * max(1.0, st.state.simTime)


@pytest.mark.parametrize("integrator", [sd.INTEGRATOR_KD, sd.INTEGRATOR_KDK])
def test_fused_pipeline_matches_reference_order(gpu, oracle, integrator):
    """The fused pipeline (deposit and reductions inside integrate, Morton
    keys) reproduces the reference-order pipeline under a fixed dt."""
    _say("fused", integrator)
    n, G, steps = 20000, 32, 8
    pos, vel = random_particles(n, seed=7, inactive=0.0)
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)  # identity carried in w
    out = []
    for pipe, key in ((sd.PIPELINE_REFERENCE, sd.KEY_ROWMAJOR), (sd.PIPELINE_FUSED, sd.KEY_MORTON)):
        s = make_sim(n, G, soft=1.5, gravity=0.5, sort_interval=4, pipeline=pipe,
                     integrator=integrator, keyMode=key)
        s.set_fixed_dt(2e-3)
        s.set_particles(pos, vel)
        s.run(steps)
        p, v = s.get_particles()
        o = np.argsort(p[:, 3])
        out.append((p[o], v[o], s.sim_time()))
        s.close()
    (p0, v0, t0), (p1, v1, t1) = out
    assert np.abs(p0[:, :3] - p1[:, :3]).max() < 1e-6
    assert np.abs(v0[:, :3] - v1[:, :3]).max() < 1e-5 * np.abs(v0[:, :3]).max()
    assert abs(t0 - t1) < 1e-9


@pytest.mark.parametrize("periodic", [False, True])
def test_fused_contact_matches_reference_order(gpu, periodic):
    """Contact steps of the fused pipeline sort the keys only: the contact
    kernel reads the particles through the permutation and the integrate
    reorders them. Against the reference-order pipeline (sort with reorder,
    contact on sorted arrays) under a fixed dt, clumps reaching the 96 cap."""
    _say("fused contact", periodic)
    n, G, steps = 30000, 32, 6
    pos, vel = random_particles(n, seed=12, inactive=0.02)
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)  # identity carried in w
    out = []
    for pipe, integ in ((sd.PIPELINE_REFERENCE, sd.INTEGRATOR_KD), (sd.PIPELINE_FUSED, sd.INTEGRATOR_KD)):
        cfg = sd.default_config(particleCount=n, gridSize=G,
                                boundary=sd.BOUNDARY_PERIODIC if periodic else sd.BOUNDARY_ISOLATED,
                                softeningCells=1.5, gravity=0.2, sortInterval=1, contactEnabled=1,
                                contactRadiusCells=0.4, contactStiffness=1e3)
        s = sd.Sim(cfg, sd.default_modes(pipeline=pipe, integrator=integ, keyMode=sd.KEY_MORTON))
        s.set_fixed_dt(1e-3)
        s.set_particles(pos, vel)
        s.run(steps)
        p, v = s.get_particles()
        live = p[:, 0] >= 0
        o = np.argsort(np.where(live, p[:, 3], np.inf))
        out.append((p[o], v[o], int(live.sum())))
        s.close()
    (p0, v0, n0), (p1, v1, n1) = out
    assert n0 == n1
    assert np.array_equal(p0[:n0, 3], p1[:n1, 3])
    assert np.abs(p0[:n0, :3] - p1[:n1, :3]).max() < 1e-6
    assert np.abs(v0[:n0, :3] - v1[:n1, :3]).max() < 1e-5 * np.abs(v0[:n0, :3]).max()


def test_fused_learned_correction_matches_reference_order(gpu):
    """Sort steps with the learned correction: the fused pipeline sorts keys
    only, the MLP reads the particles through the permutation and the
    integrate reorders; the reference-order pipeline reorders first. Same
    network (seeded weights), refreshed every step, fixed dt."""
    _say("fused learned")
    n, G, steps = 40000, 32, 6
    pos, vel = random_particles(n, seed=14, inactive=0.0)
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)  # identity carried in w
    mlp = sd.Mlp(seed=5)
    out = []
    for pipe, use_mlp in ((sd.PIPELINE_REFERENCE, True), (sd.PIPELINE_FUSED, True), (sd.PIPELINE_FUSED, False)):
        s = make_sim(n, G, soft=1.5, gravity=0.3, sort_interval=2, pipeline=pipe,
                     integrator=sd.INTEGRATOR_KD, keyMode=sd.KEY_MORTON)
        s.set_fixed_dt(1e-3)
        s.set_particles(pos, vel)
        if use_mlp:
            s.attach_mlp(mlp, 2000.0)  # seeded weights give small outputs: scale them up
            s.set_mlp_interval(1)
        s.run(steps)
        p, v = s.get_particles()
        if use_mlp:
            s.attach_mlp(None)
        o = np.argsort(p[:, 3])
        out.append((p[o], v[o]))
        s.close()
    (p0, v0), (p1, v1), (p2, v2) = out
    assert np.array_equal(p0[:, 3], p1[:, 3])
    assert np.abs(p0[:, :3] - p1[:, :3]).max() < 1e-6
    assert np.abs(v0[:, :3] - v1[:, :3]).max() < 1e-5 * np.abs(v0[:, :3]).max()
    # the correction moves the particles far beyond that tolerance (a correction
    # applied to the wrong particles would show)
    assert np.abs(v1[:, :3] - v2[:, :3]).max() > 1e-3 * np.abs(v1[:, :3]).max()


@pytest.mark.parametrize("pipeline", [sd.PIPELINE_REFERENCE, sd.PIPELINE_FUSED])
def test_contact_with_learned_correction(gpu, pipeline):
    """Contact and the learned correction together: the contact sort permutes
    the particles within the step, and each particle must still receive its
    own correction. One kick-drift step from the same state with and without
    the network differs by exactly the correction: v_a - v_b = c dt, with c
    evaluated per particle (matched through the identity in w) before the
    step. The particles start in random order, so the sort moves them."""
    _say("contact + learned", pipeline)
    n, G, dt = 30000, 32, 1e-3
    pos, vel = random_particles(n, seed=41, inactive=0.0)
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)
    mlp = sd.Mlp(seed=9)
    cfg = sd.default_config(particleCount=n, gridSize=G, softeningCells=1.5, gravity=0.2, sortInterval=1,
                            contactEnabled=1, contactRadiusCells=0.4, contactStiffness=1e3)
    out = []
    for use_mlp in (True, False):
        s = sd.Sim(cfg, sd.default_modes(pipeline=pipeline, integrator=sd.INTEGRATOR_KD, momentumZero=0))
        s.set_fixed_dt(dt)
        s.set_particles(pos, vel)
        corr = None
        if use_mlp:
            s.attach_mlp(mlp, 2000.0)
            s.set_mlp_interval(1)
            # correction at the initial positions, in the initial order; this
            # also deposits the density, so the fused pipeline then takes its
            # keys-only contact sort on the step
            corr = s.mlp_correction()
        s.step()
        p, v = s.get_particles()
        if use_mlp:
            s.attach_mlp(None)
        s.close()
        o = np.argsort(p[:, 3])
        out.append((p[o], v[o], corr))
    (pa, va, corr), (pb, vb, _) = out
    assert np.array_equal(pa[:, 3], pb[:, 3])
    dv = (va[:, :3] - vb[:, :3]).astype(np.float64)
    want = corr[:, :3].astype(np.float64) * dt
    scale = np.abs(want).max()
    assert scale > 1e-3 * np.abs(vb[:, :3]).max()  # the correction is far above rounding
    err = np.abs(dv - want).max()
    _say("  max |dv - c dt|", err, "of", scale)
    assert err < 1e-3 * scale + 4 * np.spacing(np.abs(vb[:, :3]).max())


@pytest.mark.parametrize("G,n,periodic", [(128, 2_500_000, False), (64, 2_500_000, False), (32, 300_000, False),
                                          (128, 300_000, False), (256, 2_500_000, False),
                                          (512, 2_500_000, True)])
def test_run_equals_single_steps(gpu, G, n, periodic):
    """Within a run the partials reduction of each step is carried by the next
    mesh solve's x pass (split over several blocks where the pass allows);
    single steps reduce in a launch of their own. Both sum in one fixed order,
    so a run and the same steps taken one at a time agree bit for bit, here
    under the CFL dt with momentum removal (more than 4 partials per virtual
    thread at 2.5 M particles). At S=512 (isolated G=256, periodic G=512) the
    column passes run persistent and the x pass carries the reduction."""
    _say("run vs steps", G, n, periodic)
    pos, vel = random_particles(n, seed=31)
    if periodic:
        pos[:, :3] = np.where(pos[:, 0:1] >= 0, np.mod(pos[:, :3], 1.0), pos[:, :3])
    out = []
    for single in (False, True):
        s = make_sim(n, G, periodic=periodic, gravity=0.5, sort_interval=3)
        s.set_particles(pos, vel)
        if single:
            for _ in range(7):
                s.step()
        else:
            s.run(7)
        p, v = s.get_particles()
        out.append((p, v, s.sim_time()))
        s.close()
    (p0, v0, t0), (p1, v1, t1) = out
    assert np.array_equal(p0, p1)
    assert np.array_equal(v0, v1)
    assert t0 == t1


def test_kdk_equals_shifted_kd(gpu):
    """Leapfrog from v0 equals the kick-drift scheme started from
    v0 - a(x0) dt / 2 (positions), a closed-form relation of the two."""
    n, G, steps, dt = 8000, 32, 5, 2e-3
    pos, vel = random_particles(n, seed=9, inactive=0.0)
    s = make_sim(n, G, soft=1.5, gravity=0.5, sort_interval=0, integrator=sd.INTEGRATOR_KDK)
    s.set_fixed_dt(dt)
    s.set_particles(pos, vel)
    a0 = s.particle_accel()
    s.run(steps)
    pk, vk = s.get_particles()
    s.close()
    v0s = vel.copy()
    v0s[:, :3] -= 0.5 * dt * a0[:, :3]
    s = make_sim(n, G, soft=1.5, gravity=0.5, sort_interval=0, integrator=sd.INTEGRATOR_KD,
                 momentumZero=1)
    s.set_fixed_dt(dt)
    s.set_particles(pos, v0s)
    s.run(steps)
    pd, vd = s.get_particles()
    s.close()
    assert np.abs(pk[:, :3] - pd[:, :3]).max() < 1e-6


def test_contact_vs_oracle(gpu, oracle):
    _say("contact")
    n, G = 4000, 16
    rng = np.random.default_rng(3)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.3, 0.7, (n, 3))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.01, (n, 3))
    cfg = sd.default_config(particleCount=n, gridSize=G, softeningCells=1.0, gravity=0.1,
                            sortInterval=1, contactEnabled=1, contactRadiusCells=0.5,
                            contactStiffness=1e3)
    for key in (sd.KEY_ROWMAJOR, sd.KEY_MORTON):
        s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD,
                                         keyMode=key))
        s.set_particles(pos, vel)
        s.run(3)
        gp, gv = s.get_particles()
        s.close()
        prm = oracle.make_params(G, gravity=0.1, softeningCells=1.0, sortInterval=1, keyMode=key,
                                 contact=1, contactRadiusCells=0.5, contactStiffness=1e3, allocN=n)
        st = oracle.Stepper(prm, pos, vel)
        st.step(3)
        assert np.abs(gp[:, :3] - st.pos[:, :3]).max() < 2e-6
        assert np.abs(gv[:, :3] - st.vel[:, :3]).max() < 1e-4 * np.abs(st.vel[:, :3]).max()


@pytest.mark.parametrize("key", ["rowmajor", "morton"])
def test_contact_table_reuse_vs_oracle(gpu, oracle, key):
    """Contact tables carried from one particle set to the next: a second set
    in other cells (every first-set cell now empty) must see none of the first
    set's entries. Row-major keys: the per-cell table is never cleared; Morton
    keys: the 4^3-block table is rebuilt per call and cells are searched inside
    block runs. Periodic box, sparse and dense regions, cap of 96 reached."""
    _say("contact table reuse", key)
    skey, okey = (sd.KEY_ROWMAJOR, oracle.KEY_ROWMAJOR) if key == "rowmajor" else (sd.KEY_MORTON, oracle.KEY_MORTON)
    n, G = 6000, 16
    rng = np.random.default_rng(8)
    first = np.zeros((n, 4), np.float32)
    first[:, :3] = rng.uniform(0.05, 0.45, (n, 3))
    second = np.zeros((n, 4), np.float32)
    second[: n // 2, :3] = rng.uniform(0.55, 0.75, (n // 2, 3))  # dense: cap reached
    second[n // 2:, :3] = rng.uniform(0.5, 1.0, (n - n // 2, 3))
    second[n // 2:, 0] = np.where(second[n // 2:, 0] > 0.99, 0.995, second[n // 2:, 0])  # near the wrap
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.01, (n, 3))
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC, softeningCells=1.0,
                            gravity=0.05, sortInterval=1, contactEnabled=1, contactRadiusCells=0.3,
                            contactStiffness=1e3)
    s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD,
                                     keyMode=skey))
    s.set_fixed_dt(1e-3)
    s.set_particles(first, vel)
    s.run(2)
    s.set_particles(second, vel)
    s.run(2)
    gp, gv = s.get_particles()
    s.close()
    prm = oracle.make_params(G, periodic=True, gravity=0.05, softeningCells=1.0, sortInterval=1,
                             keyMode=okey, contact=1, contactRadiusCells=0.3, contactStiffness=1e3,
                             fixedDt=1e-3, allocN=n)
    st = oracle.Stepper(prm, second, vel)
    st.step(2)
    assert np.abs(gp[:, :3] - st.pos[:, :3]).max() < 2e-6
    assert np.abs(gv[:, :3] - st.vel[:, :3]).max() < 1e-4 * np.abs(st.vel[:, :3]).max()


@pytest.mark.parametrize("key", ["rowmajor", "morton"])
def test_contact_isolated_edges_vs_oracle(gpu, oracle, key):
    """Isolated box of a side that is not a power of two (Morton codes then run
    over 128^3 codes for G = 96, taken exactly rather than rounded down to a
    power of two): particles in the edge and corner cells, whose clamped
    neighbours fold onto their own cells, and a dense corner that reaches the
    cap of 96."""
    _say("contact isolated edges", key)
    skey, okey = (sd.KEY_ROWMAJOR, oracle.KEY_ROWMAJOR) if key == "rowmajor" else (sd.KEY_MORTON, oracle.KEY_MORTON)
    n, G = 5000, 96
    rng = np.random.default_rng(17)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.002, 0.998, (n, 3))
    pos[: n // 5, :3] = rng.uniform(0.975, 0.998, (n // 5, 3))  # dense corner cells
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)  # tags the particle order
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.005, (n, 3))
    cfg = sd.default_config(particleCount=n, gridSize=G, softeningCells=1.0, gravity=0.02, sortInterval=1,
                            contactEnabled=1, contactRadiusCells=0.4, contactStiffness=1e3)
    s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD, keyMode=skey,
                                     exactGrid=1))
    s.set_fixed_dt(5e-4)
    s.set_particles(pos, vel)
    s.run(2)
    gp, gv = s.get_particles()
    s.close()
    prm = oracle.make_params(G, gravity=0.02, softeningCells=1.0, sortInterval=1, keyMode=okey, contact=1,
                             contactRadiusCells=0.4, contactStiffness=1e3, fixedDt=5e-4, allocN=n)
    st = oracle.Stepper(prm, pos, vel)
    st.step(2)
    assert np.array_equal(gp[:, 3], st.pos[:, 3])
    assert np.abs(gp[:, :3] - st.pos[:, :3]).max() < 2e-6
    assert np.abs(gv[:, :3] - st.vel[:, :3]).max() < 1e-4 * np.abs(st.vel[:, :3]).max()


def test_direct_sum_vs_fp64(gpu, oracle):
    n = 50000
    pos, vel = random_particles(n, seed=21)
    s = make_sim(n, 32)
    s.set_particles(pos, vel)
    rng = np.random.default_rng(1)
    live = np.nonzero(pos[:, 0] >= 0)[0]
    targets = rng.choice(live, 512, replace=False).astype(np.int32)
    gm, eps = 1.0 / n, 0.01
    a = s.direct_sum(targets, gm, eps)
    ref = oracle.direct_sum(pos, targets, gm, eps)
    err = np.linalg.norm(a - ref, axis=1).max() / np.linalg.norm(ref, axis=1).max()
    assert err < 2e-5
    s.close()


@pytest.mark.parametrize("G", [32, 64, 128])
def test_mass_conservation_50_steps(gpu, G):
    """sim_tests.cpp:31-52 (total grid mass stays N); exact here."""
    n = 50000
    pos, vel = random_particles(n, seed=G, inactive=0.0)
    s = make_sim(n, G, soft=3.0 * G / 128, gravity=0.6, sort_interval=40)
    s.set_particles(pos, vel)
    s.run(50)
    assert s.measure_total_grid_mass() == pytest.approx(n, abs=0)
    assert s.active_count() == n
    s.close()


def test_state_file_roundtrip(gpu, tmp_path):
    n, G = 3000, 16
    pos, vel = random_particles(n, seed=4, inactive=0.0)
    s = make_sim(n, G, integrator=sd.INTEGRATOR_KD)
    s.set_particles(pos, vel)
    s.run(3)
    p, v = s.get_particles()
    path = str(tmp_path / "state.bin")
    s.save_state(path)
    t = s.sim_time()
    s2 = sd.Sim(modes=sd.default_modes(integrator=sd.INTEGRATOR_KD))
    s2.init(sd.default_config(particleCount=10, gridSize=32))
    s2.load_state(path)
    p2, v2 = s2.get_particles()
    assert s2.grid_size() == G and s2.particle_count() == n
    assert np.array_equal(p, p2) and np.array_equal(v, v2)
    assert s2.sim_time() == pytest.approx(t, rel=1e-6)
    s.close()
    s2.close()


def test_determinism_and_measures(gpu):
    n, G = 200000, 64
    pos, vel = random_particles(n, seed=8)
    rhos = []
    for _ in range(2):
        s = make_sim(n, G, sort_interval=5)
        s.set_fixed_dt(1e-3)
        s.set_particles(pos, vel)
        s.run(12)
        rhos.append(s.get_grid(sd.GRID_DENSITY_FIXED))
        p, _ = s.get_particles()
        live = p[:, 0] >= 0
        cx, cy = s.measure_centroid()
        assert cx == pytest.approx(float(p[live, 0].astype(np.float64).mean()), abs=1e-9)
        assert s.measure_occupied_cells() == int((rhos[-1] > sd.FIX_ONE * 1e-6).sum())
        assert s.peak_cell_count() > 0
        s.close()
    assert np.array_equal(rhos[0], rhos[1])


@pytest.mark.parametrize("periodic,G", [(False, 64), (True, 64), (False, 96), (False, 128)])
def test_gather_deposit_bit_exact(gpu, oracle, periodic, G):
    """Sparse contact steps deposit in gather form (node tiles read the
    particles of the surrounding Morton block runs and store each node once).
    Where the x pass is the staged one (S != 256: the first three cases) the
    tiles store float nodes, which the next x pass reads instead of the int64
    grid; particles that move more than a cell in one step go to an int64
    side grid, and the tiles they land in are summed again. At S = 256 (the
    last case) the tiles store int64 nodes and such particles are added by an
    atomic pass. Two runs step in lockstep: A solves its potential from the
    step's own deposit, B from the int64 grid of a fresh atomic deposit (its
    density is read first). The potentials agree bit for bit and both
    densities equal the oracle's fixed-point deposit of the new positions, on
    the first step (sort with reorder) and later ones (keys-only sort, reorder
    in the integrate)."""
    _say("gather deposit", periodic, G)
    n, dt = 30000, 1e-3
    rng = np.random.default_rng(61)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.003, 0.997, (n, 3))
    pos[: n // 10, :3] = rng.choice([0.003, 0.997], (n // 10, 3))  # edge and corner cells
    pos[:, 3] = np.arange(1, n + 1)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.2 / G / dt, (n, 3))       # ~0.2 cells per step
    fast = rng.random(n) < 0.03
    vel[fast, :3] = rng.choice([-1.0, 1.0], (int(fast.sum()), 3)) * (2.5 / G / dt)  # 2.5 cells per step
    cfg = sd.default_config(particleCount=n, gridSize=G,
                            boundary=sd.BOUNDARY_PERIODIC if periodic else sd.BOUNDARY_ISOLATED,
                            softeningCells=1.0, gravity=0.05, sortInterval=1, contactEnabled=1,
                            contactRadiusCells=0.2, contactStiffness=1e3)
    sims = [sd.Sim(cfg, sd.default_modes(integrator=sd.INTEGRATOR_KD, exactGrid=1)) for _ in range(2)]
    for s in sims:
        s.set_fixed_dt(dt)
        s.set_particles(pos, vel)
    a, b = sims
    for step in range(3):
        a.step()
        b.step()
        pot_a = a.get_grid(sd.GRID_POTENTIAL)
        rho_a = a.get_grid(sd.GRID_DENSITY_FIXED)
        rho_b = b.get_grid(sd.GRID_DENSITY_FIXED)
        pot_b = b.get_grid(sd.GRID_POTENTIAL)
        p, _ = a.get_particles()
        ref = oracle.deposit_fixed(p, G, periodic)
        _say("  step", step, "mass", int(rho_a.sum()) // sd.FIX_ONE)
        assert np.array_equal(rho_a, ref)
        assert np.array_equal(rho_b, ref)
        assert np.array_equal(pot_a.view(np.uint32), pot_b.view(np.uint32))
    for s in sims:
        s.close()
