"""The reference's step policies and formats on the GPU, against the oracle
and closed forms:

* the integrator's boundary and force terms (Sim.cu:685-906): soft-sphere
  boundary, dark energy (with the soft-sphere taper), the light-speed clamp
  (lightSpeedFor, Sim.h:131-148) and the cube walls, plus the time scale of
  the CFL policy (dtScaleFor, Sim.h:89-93) -- reference kernel order and the
  fused pipeline against the oracle's Sim::step;
* the harmonic oscillator of the dark-energy term with gravity off, whose
  kick-drift steps are a closed-form matrix power;
* STARDUST v1/v2 state files built byte by byte per Sim.cu:4980-5120, and the
  file this library writes parsed the same way;
* peakCellCount (Sim.h:912-915) against a count of the cell keys.
"""
import struct
import sys

import numpy as np
import pytest

import stardust_amd as sd

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _light_c2(length_scale, time_unit_scale):
    """c^2 of lightSpeedFor (Sim.h:131-148) as the library computes it."""
    L = 1.0e5 * max(length_scale, 1e-6)
    T = 1.0e7 * max(time_unit_scale, 1e-6)
    c = np.float32(T / L)
    return float(np.float32(c * c))


def _particles(n, seed, edge=False):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.05, 0.95, (n, 3))
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.05, (n, 3))
    if edge:
        # a quarter just inside the cube walls, moving out
        m = n // 4
        side = rng.integers(0, 2, (m, 3))
        pos[:m, :3] = np.where(side == 1, 0.9975, 0.0025)
        vel[:m, :3] = np.where(side == 1, 1.0, -1.0) * rng.uniform(0.5, 2.0, (m, 3))
    pos[:, 3] = np.arange(1, n + 1, dtype=np.float32)  # identity in w
    return pos, vel


CASES = {
    # name: (config fields, oracle fields, fixed dt (0 = CFL policy), edge particles)
    "soft_sphere": (dict(softBoundR=0.3), dict(softBoundR=0.3), 2e-3, False),
    "dark_energy": (dict(darkEnergy=0.8, softBoundR=0.35), dict(darkEnergy=0.8, softBoundR=0.35), 2e-3, False),
    "light_clamp": (dict(timeUnitScale=4e-4), dict(c2=_light_c2(1.0, 4e-4)), 2e-3, False),
    "walls": (dict(), dict(), 2e-3, True),
    "time_scale_0.25": (dict(timeScale=0.25), dict(timeScale=0.25), 0.0, False),
    "time_scale_6": (dict(timeScale=6.0), dict(timeScale=6.0), 0.0, False),
}


@pytest.mark.parametrize("case", list(CASES))
def test_step_policies_vs_oracle(gpu, oracle, case):
    cfg_kw, orc_kw, dt, edge = CASES[case]
    n, G, steps = 20000, 32, 6
    pos, vel = _particles(n, 51, edge)
    prm = oracle.make_params(G, gravity=0.4, softeningCells=1.5, sortInterval=3, fixedDt=dt, allocN=n,
                             keyMode=oracle.KEY_ROWMAJOR, **orc_kw)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    o = np.argsort(st.pos[:, 3])
    op, ov = st.pos[o], st.vel[o]
    if case == "light_clamp":
        # the clamp binds (momentum removal after it shifts speeds by the mean)
        sp = np.linalg.norm(ov[:, :3].astype(np.float64), axis=1)
        assert sp.max() <= np.sqrt(prm.c2) * 1.01 and (sp > 0.99 * np.sqrt(prm.c2)).sum() > 100
    pipes = [(sd.PIPELINE_REFERENCE, sd.KEY_ROWMAJOR)]
    if dt > 0:  # the fused pipeline takes vmax before the mean removal: compared under a fixed dt
        pipes.append((sd.PIPELINE_FUSED, sd.KEY_MORTON))
    for pipe, key in pipes:
        _say("policy", case, "pipeline", pipe)
        cfg = sd.default_config(particleCount=n, gridSize=G, gravity=0.4, softeningCells=1.5, sortInterval=3,
                                **cfg_kw)
        s = sd.Sim(cfg, sd.default_modes(pipeline=pipe, integrator=sd.INTEGRATOR_KD, keyMode=key))
        s.set_fixed_dt(dt)
        s.set_particles(pos, vel)
        s.run(steps)
        gp, gv = s.get_particles()
        t = s.sim_time()
        s.close()
        g = np.argsort(gp[:, 3])
        gp, gv = gp[g], gv[g]
        assert np.array_equal(gp[:, 3], op[:, 3])
        dp = np.abs(gp[:, :3] - op[:, :3]).max()
        dv = np.abs(gv[:, :3] - ov[:, :3]).max() / np.abs(ov[:, :3]).max()
        _say("  max |dx|", dp, "max rel dv", dv, "t", t, st.state.simTime)
        assert dp < 2e-6
        assert dv < 1e-4
        assert abs(t - st.state.simTime) < 1e-6 * st.state.simTime


@pytest.mark.parametrize("pipeline", [sd.PIPELINE_REFERENCE, sd.PIPELINE_FUSED])
def test_harmonic_dark_energy_closed_form(gpu, pipeline):
    """Gravity off and dark energy -w^2: each particle oscillates about the
    box centre, and k kick-drift steps are [y, v] <- M^k [y, v] with
    M = [[1 - w^2 dt^2, dt], [-w^2 dt, 1]] (Sim.cu:822, 842), across sorts."""
    n, k, dt, w2 = 50000, 12, 2e-3, 40.0
    rng = np.random.default_rng(52)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.4, 0.6, (n, 3))
    pos[:, 3] = np.arange(1, n + 1)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.05, (n, 3))
    cfg = sd.default_config(particleCount=n, gridSize=32, gravity=0.0, softeningCells=1.0, sortInterval=4,
                            darkEnergy=-w2)
    s = sd.Sim(cfg, sd.default_modes(pipeline=pipeline, integrator=sd.INTEGRATOR_KD, momentumZero=0))
    s.set_fixed_dt(dt)
    s.set_particles(pos, vel)
    s.run(k)
    gp, gv = s.get_particles()
    s.close()
    o = np.argsort(gp[:, 3])
    M = np.array([[1.0 - w2 * dt * dt, dt], [-w2 * dt, 1.0]])
    Mk = np.linalg.matrix_power(M, k)
    y0, v0 = pos[:, :3].astype(np.float64) - 0.5, vel[:, :3].astype(np.float64)
    y = Mk[0, 0] * y0 + Mk[0, 1] * v0
    v = Mk[1, 0] * y0 + Mk[1, 1] * v0
    dy = np.abs(gp[o, :3] - 0.5 - y).max()
    dv = np.abs(gv[o, :3] - v).max()
    _say("harmonic", pipeline, "max |dy|", dy, "max |dv|", dv)
    assert dy < 1e-6  # FP32 positions near 0.5 over 12 steps; a kick-drift order or sign mistake gives 1.6e-5 per step
    assert dv < 3e-6


HEADER = "<8siiiiiff4fi5i"  # Sim.cu:4980-4994: 76 bytes


def _write_state(path, version, pos, vel, G, boundary, gravity, sim_time, preset=0):
    n = pos.shape[0]
    h = struct.pack(HEADER, b"STARDUST", version, n, G, preset, boundary, gravity, sim_time, 0.0, 0.0, 0.0, 0.0, 0,
                    0, 0, 0, 0, 0)
    assert len(h) == 76
    with open(path, "wb") as f:
        f.write(h)
        f.write(np.ascontiguousarray(pos, np.float32).tobytes())
        f.write(np.ascontiguousarray(vel, np.float32).tobytes())


@pytest.mark.parametrize("version", [1, 2])
def test_state_file_layout(gpu, tmp_path, version):
    """A STARDUST file written byte by byte per Sim.cu:4980-5120 loads into the
    library; the file the library saves parses back the same way."""
    rng = np.random.default_rng(version)
    n, G = 5000, 32
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.1, 0.9, (n, 3))
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.02, (n, 3))
    path = str(tmp_path / "in.bin")
    _write_state(path, version, pos, vel, G, 1, 0.75, 1.25, preset=0)
    s = sd.Sim(sd.default_config(particleCount=10, gridSize=16), sd.default_modes(integrator=sd.INTEGRATOR_KD))
    s.load_state(path)
    assert s.grid_size() == G and s.particle_count() == n
    assert s.config().boundary == 1 and s.config().gravity == pytest.approx(0.75)
    assert s.sim_time() == pytest.approx(1.25)
    p, v = s.get_particles()
    assert np.array_equal(p, pos) and np.array_equal(v, vel)
    s.run(2)
    out = str(tmp_path / "out.bin")
    s.save_state(out)
    p2, v2 = s.get_particles()
    t2 = s.sim_time()
    s.close()
    raw = open(out, "rb").read()
    f = struct.unpack(HEADER, raw[:76])
    assert f[0] == b"STARDUST" and f[1] == 2 and f[2] == n and f[3] == G and f[5] == 1
    assert f[6] == pytest.approx(0.75) and f[7] == pytest.approx(t2, rel=1e-6)
    assert f[12] == 0 and f[13] == 0  # no black holes
    assert len(raw) == 76 + 2 * 16 * n
    fp = np.frombuffer(raw[76:76 + 16 * n], np.float32).reshape(n, 4)
    fv = np.frombuffer(raw[76 + 16 * n:], np.float32).reshape(n, 4)
    assert np.array_equal(fp, p2) and np.array_equal(fv, v2)


def test_state_file_refusals(gpu, tmp_path):
    s = sd.Sim(sd.default_config(particleCount=10, gridSize=16))
    pos = np.full((4, 4), 0.5, np.float32)
    bad = str(tmp_path / "bad.bin")
    _write_state(bad, 2, pos, pos, 16, 0, 1.0, 0.0)
    raw = bytearray(open(bad, "rb").read())
    raw[0:8] = b"STARDUSX"
    open(bad, "wb").write(raw)
    with pytest.raises(sd.StardustError):
        s.load_state(bad)
    bh = str(tmp_path / "bh.bin")
    h = struct.pack(HEADER, b"STARDUST", 2, 4, 16, 0, 0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 1, 0, 0, 0, 0)
    open(bh, "wb").write(h + bytes(64))
    with pytest.raises(sd.StardustError):
        s.load_state(bh)  # black holes are off this path
    assert not sd.failed()
    s.close()


def test_peak_cell_count(gpu, oracle):
    """peakCellCount: the most particles in one cell of the key convention
    (int)(p G); it allocates nothing (it borrows the acceleration grid, which a
    later read rebuilds from the potential)."""
    n, G = 200000, 32
    rng = np.random.default_rng(53)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.clip(rng.normal(0.5, 0.08, (n, 3)), 0.01, 0.99)
    pos[::97] = (-1, -1, -1, 0)
    s = sd.Sim(sd.default_config(particleCount=n, gridSize=G))
    s.set_particles(pos, np.zeros_like(pos))
    s.compute_forces()
    acc0 = s.get_grid(sd.GRID_ACCEL)
    keys, _ = oracle.cell_keys(pos, G, oracle.KEY_ROWMAJOR)
    want = int(np.bincount(keys[keys > 0]).max())
    assert s.peak_cell_count() == want
    assert np.array_equal(s.get_grid(sd.GRID_ACCEL), acc0)
    s.close()


def test_density_device_views(gpu):
    """Sim::densityDevicePtr / fieldDevicePtr(Density) (Sim.cu:5151, 5229-5240):
    the G^3 density of the current positions and its sum along z, read
    straight from device memory; the off-path fields are refused."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n, G = 100000, 32
    rng = np.random.default_rng(54)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.05, 0.95, (n, 3))
    s = sd.Sim(sd.default_config(particleCount=n, gridSize=G, sortInterval=2))
    s.set_particles(pos, np.zeros_like(pos))
    s.run(3)
    rho = s.get_grid(sd.GRID_DENSITY)
    d = np.empty((G, G, G), np.float32)
    assert hip.hipMemcpy(d.ctypes.data, s.density_device_ptr(), d.nbytes, 2) == 0
    assert np.array_equal(d, rho)
    proj = np.empty((G, G), np.float32)
    assert hip.hipMemcpy(proj.ctypes.data, s.field_device_ptr(sd.FIELD_DENSITY), proj.nbytes, 2) == 0
    want = np.zeros((G, G), np.float32)
    for z in range(G):
        want += rho[z]
    assert np.array_equal(proj, want)
    assert abs(float(proj.astype(np.float64).sum()) - n) < 1e-2
    with pytest.raises(sd.StardustError):
        s.field_device_ptr(2)  # light: star physics, off this path
    assert not sd.failed()
    # the views do not disturb the step: a run continues as before
    s.run(2)
    assert s.measure_total_grid_mass() == pytest.approx(n, abs=0)
    s.close()


def _gas_particles(n, seed):
    """Clumped particles (several per cell): 70% gas (pos.w 0), 20% stars
    (pos.w 1), 10% dark matter (vel.w -100)."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    centres = rng.uniform(0.3, 0.7, (6, 3))
    pos[:, :3] = centres[rng.integers(0, 6, n)] + rng.normal(0, 0.03, (n, 3))
    pos[:, :3] = np.clip(pos[:, :3], 0.01, 0.99)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.05, (n, 3))
    kind = rng.random(n)
    pos[(kind >= 0.7) & (kind < 0.9), 3] = 1.0
    vel[kind >= 0.9, 3] = -100.0
    return pos, vel


@pytest.mark.parametrize("contact,periodic,key", [(False, False, "rowmajor"), (False, True, "morton"),
                                                  (True, False, "morton"), (True, True, "rowmajor")])
def test_gas_physics_vs_oracle(gpu, oracle, contact, periodic, key):
    """Pressure and cooling (Sim.cu:1212-1415, 3605-3633, 3732-3830) in the
    reference kernel order: the cooling pass runs every 2 steps (every step
    with contact), sums each cell's velocities in sorted order and pulls gas
    toward the cell mean; the pressure grid from its dispersion sums acts on
    gas. Against the oracle's Sim::step with the same key order."""
    n, G, steps, dt = 20000, 32, 6, 2e-3
    pos, vel = _gas_particles(n, 81)
    if periodic:
        pos[:, :3] = np.mod(pos[:, :3], 1.0)
    skey, okey = (sd.KEY_ROWMAJOR, oracle.KEY_ROWMAJOR) if key == "rowmajor" else (sd.KEY_MORTON, oracle.KEY_MORTON)
    kw = dict(pressure=1, cooling=1, pressureK=0.45, coolingRate=0.9)
    prm = oracle.make_params(G, periodic=periodic, gravity=0.3, softeningCells=1.5, sortInterval=3, keyMode=okey,
                             fixedDt=dt, contact=int(contact), contactRadiusCells=0.3, contactStiffness=1e3,
                             allocN=n, **kw)
    st = oracle.Stepper(prm, pos, vel)
    st.step(steps)
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=int(periodic), gravity=0.3, softeningCells=1.5,
                            sortInterval=3, contactEnabled=int(contact), contactRadiusCells=0.3,
                            contactStiffness=1e3, pressureEnabled=1, coolingEnabled=1, pressureK=0.45,
                            coolingRate=0.9)
    out = {}
    for gas in (True, False):
        c = cfg
        if not gas:
            c = sd.default_config(**{f: getattr(cfg, f) for f, _ in sd.Config._fields_})
            c.pressureEnabled = 0
            c.coolingEnabled = 0
        s = sd.Sim(c, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD, keyMode=skey))
        s.set_fixed_dt(dt)
        s.set_particles(pos, vel)
        s.run(steps)
        out[gas] = s.get_particles()
        s.close()
    gp, gv = out[True]
    _say("gas", contact, periodic, key)
    assert np.array_equal(gp[:, 0] < 0, st.pos[:, 0] < 0)
    assert np.array_equal(gp[:, 3], st.pos[:, 3]) and np.array_equal(gv[:, 3], st.vel[:, 3])  # same order
    dp = np.abs(gp[:, :3] - st.pos[:, :3]).max()
    dv = np.abs(gv[:, :3] - st.vel[:, :3]).max() / np.abs(st.vel[:, :3]).max()
    _say("  max |dx|", dp, "max rel dv", dv)
    assert dp < 2e-6
    assert dv < 1e-4
    # the gas terms change the motion far beyond that tolerance
    pv, vv = out[False]
    assert np.abs(vv[:, :3] - gv[:, :3]).max() > 1e-2 * np.abs(gv[:, :3]).max()


def test_gas_physics_needs_reference_order(gpu):
    cfg = sd.default_config(particleCount=1000, gridSize=16, pressureEnabled=1)
    with pytest.raises(sd.StardustError):
        sd.Sim(cfg, sd.default_modes())  # fused pipeline, leapfrog
    s = sd.Sim(cfg, sd.default_modes(pipeline=sd.PIPELINE_REFERENCE, integrator=sd.INTEGRATOR_KD))
    with pytest.raises(sd.StardustError):
        s.set_modes(sd.default_modes())
    assert not sd.failed()
    s.close()
