"""GPU tests of the learned correction: BF16 MFMA forward / backward against
the FP64 oracle, the fused per-particle inference against the feature +
forward path, and training on the GPU. Tolerances are those of BF16 weights
and activations (8-bit mantissa) with FP32 accumulation."""
import sys

import numpy as np
import pytest

import stardust_amd as sd

pytestmark = pytest.mark.gpu
H = 128


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    nW = 30 * H + 2 * H * H + 3 * H
    fan = np.concatenate([np.full(30 * H, 30), np.full(2 * H * H, H), np.full(3 * H, H)])
    W = (rng.uniform(-1, 1, nW) * np.sqrt(6.0 / fan) * scale).astype(np.float32)
    b = (rng.normal(0, 0.1, 3 * H + 3)).astype(np.float32)
    return W, b


def bf16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def test_mlp_forward_vs_oracle(gpu, oracle):
    _say("mlp forward")
    W, b = params(1)
    m = sd.Mlp(seed=1)
    m.set_params(W, b)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (1000, 30)).astype(np.float32)
    Y = m.forward(X)
    # FP64 oracle on the BF16-rounded weights and inputs
    Yr = oracle.mlp_forward(bf16_round(W).astype(np.float64), b.astype(np.float64), H,
                            bf16_round(X).astype(np.float64))
    err = np.abs(Y - Yr).max() / np.abs(Yr).max()
    _say("  rel max err", err)
    assert err < 3e-2
    assert np.corrcoef(Y.ravel(), Yr.ravel())[0, 1] > 0.999


def test_mlp_gradients_vs_oracle(gpu, oracle):
    _say("mlp grad")
    W, b = params(3)
    m = sd.Mlp(seed=1)
    m.set_params(W, b)
    rng = np.random.default_rng(4)
    n = 2048
    X = rng.normal(0, 1, (n, 30)).astype(np.float32)
    T = rng.normal(0, 1, (n, 3)).astype(np.float32)
    loss, gW, gb = m.grad(X, T)
    lr, gWr, gbr = oracle.mlp_loss_grad(bf16_round(W).astype(np.float64), b.astype(np.float64), H,
                                        bf16_round(X).astype(np.float64), T.astype(np.float64))
    assert loss == pytest.approx(lr, rel=2e-2)
    offs = [0, 30 * H, 30 * H + H * H, 30 * H + 2 * H * H, 30 * H + 2 * H * H + 3 * H]
    for l in range(4):
        g, r = gW[offs[l]:offs[l + 1]], gWr[offs[l]:offs[l + 1]]
        cos = float(np.dot(g, r) / (np.linalg.norm(g) * np.linalg.norm(r)))
        rel = float(np.linalg.norm(g - r) / np.linalg.norm(r))
        _say(f"  layer {l}: cos {cos:.5f} rel {rel:.4f}")
        assert cos > 0.995 and rel < 0.08
    rel_b = float(np.linalg.norm(gb - gbr) / np.linalg.norm(gbr))
    assert rel_b < 0.08


def test_mlp_training_reduces_loss(gpu):
    _say("mlp train")
    rng = np.random.default_rng(5)
    n = 32768
    X = rng.normal(0, 1, (n, 30)).astype(np.float32)
    A = rng.normal(0, 0.3, (30, 3))
    T = (np.tanh(X @ A) + 0.1 * X[:, :3]).astype(np.float32)
    m = sd.Mlp(seed=7)
    l0, _, _ = m.grad(X[:8192], T[:8192])
    losses = m.train(X, T, epochs=8, batch=4096, lr=2e-3, seed=3)
    l1, _, _ = m.grad(X[:8192], T[:8192])
    _say("  loss", l0, "->", l1, losses)
    assert l1 < 0.2 * l0
    assert np.all(np.isfinite(losses))


def test_features_and_fused_inference(gpu, oracle):
    """Features match a numpy restatement on the oracle density; the fused
    per-particle kernel matches features -> forward -> scale."""
    _say("mlp features/infer")
    n, G = 20000, 32
    rng = np.random.default_rng(6)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.1, 0.9, (n, 3))
    pos[::50] = (-1, -1, -1, 0)
    vel = np.zeros_like(pos)
    gravity = 0.7
    s = sd.Sim(sd.default_config(particleCount=n, gridSize=G, gravity=gravity, softeningCells=1.0),
               sd.default_modes())
    s.set_particles(pos, vel)
    targets = np.arange(0, n, 7, dtype=np.int32)
    targets = targets[pos[targets, 0] >= 0]
    X = s.features(targets)
    rho = oracle.deposit_fixed(pos, G).astype(np.float64) / oracle.FIX_ONE
    p = pos[targets, :3]
    c = np.clip(np.floor(p * np.float32(G)).astype(int), 0, G - 1)
    ref = np.zeros((targets.size, 30))
    q = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x = np.clip(c[:, 0] + dx, 0, G - 1)
                y = np.clip(c[:, 1] + dy, 0, G - 1)
                z = np.clip(c[:, 2] + dz, 0, G - 1)
                ref[:, q] = np.log1p(rho[z, y, x])
                q += 1
    ref[:, 27:] = 2.0 * (p.astype(np.float64) * G - 0.5 - c)
    assert np.abs(X - ref).max() < 1e-5
    m = sd.Mlp(seed=11)
    W, b = params(12, scale=0.5)
    m.set_params(W, b)
    out_scale = 0.37
    s.attach_mlp(m, out_scale)
    corr = s.mlp_correction()
    a_unit = np.float32(gravity) * G / n
    Y = m.forward(X) * out_scale * a_unit
    err = np.abs(corr[targets, :3] - Y).max() / np.abs(Y).max()
    _say("  fused vs path rel err", err)
    assert err < 2e-2
    assert np.all(corr[pos[:, 0] < 0] == 0)
    # an attached network changes the step (and detaching restores it)
    s.set_fixed_dt(1e-3)
    s.run(2)
    s.attach_mlp(None)
    s.close()


def test_fused_inference_vs_oracle(gpu, oracle):
    """The fused inference kernel (FP16 operands: the BF16 weights converted
    exactly, FP16 activations, polynomial GELU on packed FP16) against the
    FP64 oracle forward on the same BF16 weights and FP64 features, beside the
    BF16 GEMM path (BF16 activations, FP32 tanh GELU): the FP16 kernel computes
    the BF16 network at least as closely as the BF16-activation path does. The
    packaged trained network on a disk of 200,000 particles, G=128."""
    from stardust_amd import ic, learned
    _say("fused inference vs oracle")
    d = ic.disk_scene(200000, 128, 5)
    pos, vel = d["pos"], d["vel"]
    n, G = pos.shape[0], 128
    gravity = d["gravity"]
    s = sd.Sim(sd.default_config(particleCount=n, gridSize=G, gravity=gravity, softeningCells=1.0),
               sd.default_modes())
    s.set_particles(pos, vel)
    m, out_scale = learned.load_model()
    W, b = m.get_params()
    targets = np.sort(np.random.default_rng(3).choice(n, 4096, replace=False)).astype(np.int32)
    # FP64 features from the oracle's density
    rho = oracle.deposit_fixed(pos, G).astype(np.float64) / oracle.FIX_ONE
    p = pos[targets, :3]
    c = np.clip(np.floor(p * np.float32(G)).astype(int), 0, G - 1)
    X = np.zeros((targets.size, 30))
    q = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                X[:, q] = np.log1p(rho[np.clip(c[:, 2] + dz, 0, G - 1), np.clip(c[:, 1] + dy, 0, G - 1),
                                       np.clip(c[:, 0] + dx, 0, G - 1)])
                q += 1
    X[:, 27:] = 2.0 * (p.astype(np.float64) * G - 0.5 - c)
    a_unit = float(np.float32(gravity) * G / n)
    ref = oracle.mlp_forward(bf16_round(W).astype(np.float64), b.astype(np.float64), H, X) * out_scale * a_unit
    s.attach_mlp(m, out_scale)
    fused = s.mlp_correction()[targets, :3].astype(np.float64)
    s.attach_mlp(None)
    path = m.forward(s.features(targets)).astype(np.float64) * out_scale * a_unit
    s.close()

    def rel(a):
        return float(np.sqrt(((a - ref) ** 2).sum() / (ref ** 2).sum()))
    e_fused, e_path = rel(fused), rel(path)
    _say("  rel rms vs oracle: fused FP16", e_fused, "BF16 path", e_path)
    assert e_fused < 5e-3
    assert e_fused <= e_path


def test_correction_on_sparse_contact_steps(gpu):
    """Contact steps at under half a particle per cell deposit in gather form,
    with float nodes where the x pass is the staged one (periodic G=64: S=64).
    A learned correction recomputed on such a step deposits the int64 grid it
    reads afresh, while the mesh, forked on the second stream before that
    deposit, keeps reading the float nodes. Run A just steps; run B reads its
    int64 density after every step, so its next correction and mesh read that
    grid instead. Both runs stay bit-identical."""
    _say("mlp on sparse contact steps")
    n, G, dt = 20000, 64, 1e-3
    rng = np.random.default_rng(71)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, 1.0, (n, 3))
    pos[:, 3] = np.arange(1, n + 1)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.3 / G / dt, (n, 3))
    cfg = sd.default_config(particleCount=n, gridSize=G, boundary=sd.BOUNDARY_PERIODIC, softeningCells=1.0,
                            gravity=0.05, sortInterval=1, contactEnabled=1, contactRadiusCells=0.2,
                            contactStiffness=1e3)
    m = sd.Mlp(seed=11)
    W, b = params(12, scale=0.5)
    m.set_params(W, b)
    sims = [sd.Sim(cfg, sd.default_modes(integrator=sd.INTEGRATOR_KD, exactGrid=1)) for _ in range(2)]
    for s in sims:
        s.set_fixed_dt(dt)
        s.set_particles(pos, vel)
        s.attach_mlp(m, 0.37)
        s.set_mlp_interval(1)
    a, b = sims
    for step in range(4):
        a.step()
        b.step()
        b.get_grid(sd.GRID_DENSITY_FIXED)
    pa, va = a.get_particles()
    pb, vb = b.get_particles()
    _say("  max |v|", float(np.abs(va[:, :3]).max()))
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    for s in sims:
        s.close()
