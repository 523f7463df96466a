"""The C-ABI library loads and exports every symbol declared in
include/stardust/sd_api.h (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

import stardust_amd as sd

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "stardust", "sd_api.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(sd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 50
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []


def test_python_mirror_declares_every_symbol():
    sd.lib()
    import inspect
    src = inspect.getsource(sd._declare)
    missing = [s for s in declared_symbols() if f'"{s}"' not in src]
    assert missing == []


def test_defaults_match_simconfig():
    """sd_config_default carries SimConfig's defaults (Sim.h:157-796)."""
    c = sd.Config()
    sd.lib().sd_config_default(C.byref(c))
    assert (c.particleCount, c.gridSize, c.sortInterval) == (2000000, 128, 40)
    assert c.gravity == pytest.approx(0.6) and c.softeningCells == pytest.approx(3.0)
    assert c.contactStiffness == pytest.approx(1e6) and c.contactDamping == pytest.approx(0.35)
    assert c.boundary == sd.BOUNDARY_ISOLATED and c.preset == sd.PRESET_FILAMENT
    assert c.pressureEnabled == 1 and c.coolingEnabled == 1
    m = sd.default_modes()
    assert (m.keyMode, m.depositMode, m.integrator, m.pipeline, m.momentumZero) == (
        sd.KEY_MORTON, sd.DEPOSIT_FIXED, sd.INTEGRATOR_KDK, sd.PIPELINE_FUSED, 1)


def test_statics_without_device():
    L = sd.lib()
    assert L.sd_max_grid_size(sd.BOUNDARY_ISOLATED) == 256
    assert L.sd_max_grid_size(sd.BOUNDARY_PERIODIC) == 512
    b1 = L.sd_estimate_bytes(1000000, 128, 0)
    b2 = L.sd_estimate_bytes(2000000, 128, 0)
    assert b2 > b1 > 0
    assert L.sd_max_particles_for(128, 0, 0) == 0
    assert L.sd_max_particles_for(128, 0, 1 << 40) > 1000000
    assert L.sd_device_available() in (0, 1)


def test_config_denials_are_not_sticky():
    """Off-path physics and the Filament preset are refused with SD_ERR_ARG."""
    s = sd.lib().sd_create()
    try:
        c = sd.Config()
        sd.lib().sd_config_default(C.byref(c))  # reference defaults: pressure/cooling on
        assert sd.lib().sd_init(s, C.byref(c)) == sd.SD_ERR_ARG
        c2 = sd.default_config(blackHoleEnabled=1)
        assert sd.lib().sd_init(s, C.byref(c2)) == sd.SD_ERR_ARG
        assert sd.lib().sd_failed() == 0
    finally:
        sd.lib().sd_destroy(s)
