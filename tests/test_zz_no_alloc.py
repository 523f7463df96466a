"""The allocation-free direct sum (sd_direct_sum_noalloc) against
sd_direct_sum through the C ABI: device-resident targets and outputs, FP64
partials in the library's fixed per-device scratch. Bit-identical sums on two
calls in a row, the second with a smaller input (fewer particles and
targets), the first large enough that its targets run in several groups."""
import ctypes as C
import sys

import numpy as np
import pytest

import stardust_amd as sd

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers for the test's own inputs and outputs (the caller's side)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _particles(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.05, 0.95, (n, 3))
    pos[rng.random(n) < 0.03] = (-1, -1, -1, 0)  # inactive slots
    vel = np.zeros_like(pos)
    return pos, vel


def test_direct_sum_noalloc_matches_direct_sum(gpu):
    L = sd.lib()
    dev = _Dev()
    try:
        # n = 1,000,000 gives 62 source chunks: 11,275 targets per scratch
        # group, so 20,000 targets take two groups; then a smaller call
        # n = 1,000,000 gives 62 source chunks, so 11,275 targets per scratch
        # group: 20,000 targets take two groups; then smaller calls, on the
        # same handle and on a handle with fewer particles
        for n, nts, seed in ((1_000_000, (20_000, 500), 1), (50_000, (777,), 2)):
            pos, vel = _particles(n, seed)
            s = sd.Sim(sd.default_config(particleCount=n, gridSize=32))
            s.set_particles(pos, vel)
            gm, eps = 1.0 / n, 0.004
            for nt in nts:
                print("[case] direct sum noalloc", n, nt, flush=True)
                sys.stdout.flush()
                targets = np.random.default_rng(seed + nt).choice(n, nt, replace=False).astype(np.int32)
                want = s.direct_sum(targets, gm, eps)
                tdev = dev.upload(targets)
                odev = dev.alloc(want.nbytes)
                assert L.sd_direct_sum_noalloc(s._h, tdev, nt, gm, eps, odev) == sd.SD_OK
                s.synchronize()
                got = dev.download(odev, want)
                assert np.all(np.isfinite(got))
                assert np.array_equal(got, want)
                # argument errors are returned, not sticky
                assert L.sd_direct_sum_noalloc(s._h, tdev, 0, gm, eps, odev) == sd.SD_ERR_ARG
                assert L.sd_direct_sum_noalloc(s._h, None, nt, gm, eps, odev) == sd.SD_ERR_ARG
                assert not L.sd_failed()
            s.close()
    finally:
        dev.free()
