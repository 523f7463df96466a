"""The learned correction trained on the GPU reduces the coarse-mesh force
error against the direct sum on a held-out scene (reduced sizes)."""
import numpy as np
import pytest

from stardust_amd import learned

pytestmark = pytest.mark.gpu


def test_trained_correction_beats_plain_coarse_mesh(gpu):
    scenes = [("disk", 200000, 11, 12.0), ("pair", 120000, 12, 20.0), ("disk", 300000, 13, 16.0),
              ("pair", 150000, 14, 22.0)]
    X, T = learned.training_pairs(G=64, scenes=scenes, n_targets=16384, evolve=(0, 10),
                                  log=lambda *a: None)
    assert X.shape == (8 * 16384, 30) and np.all(np.isfinite(T))
    mlp, out_scale, losses = learned.train(X, T, epochs=20, batch=4096, lr=2e-3, log=lambda *a: None)
    assert losses[-1] < losses[0]
    p4, v4, M, L = learned.scene_pair(150000, seed=99, L=20.0)
    err = learned.evaluate(p4, v4, M, L, 64, 128, mlp, out_scale, n_targets=8192)
    print(err)
    assert err["coarse+mlp"]["rel_rms"] < err["coarse"]["rel_rms"]
    assert err["fine"]["rel_rms"] < err["coarse"]["rel_rms"]
