"""stardust_amd -- MI355X particle-mesh gravity step behind Stardust's Sim API.

The compute path is ``libstardust_amd.so`` (hand-written gfx950 HIP kernels
behind the C ABI in ``include/stardust/sd_api.h``). This module is a thin
ctypes host mirror of the reference's ``class Sim`` (src/sim/Sim.h:854-1135)
for Python callers and tests; it never computes physics itself and raises if
the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstardust_amd.so")
HEADER_PATH = os.path.join(_HERE, "..", "include", "stardust", "sd_api.h")

# status codes / enums (sd_api.h)
SD_OK = 0
SD_ERR_FAILED = -1
SD_ERR_ARG = -2
SD_ERR_DEVICE = -3
SD_ERR_STATE = -4
SD_ERR_IO = -5
BOUNDARY_ISOLATED = 0
BOUNDARY_PERIODIC = 1
PRESET_FILAMENT = 0
PRESET_EMPTY = 1
KEY_ROWMAJOR = 0
KEY_MORTON = 1
DEPOSIT_FIXED = 0
DEPOSIT_FLOAT = 1
INTEGRATOR_KD = 0
INTEGRATOR_KDK = 1
PIPELINE_REFERENCE = 0
PIPELINE_FUSED = 1
GRID_DENSITY_FIXED = 0
GRID_DENSITY = 1
FIELD_DENSITY = 0
GRID_POTENTIAL = 2
GRID_ACCEL = 3
FIX_ONE = 1 << 24


class StardustError(RuntimeError):
    pass


class Config(C.Structure):
    """sd_config: the SimConfig fields of the PM path (Sim.h:157-796)."""

    _fields_ = [
        ("particleCount", C.c_int),
        ("gridSize", C.c_int),
        ("boundary", C.c_int),
        ("preset", C.c_int),
        ("gravity", C.c_float),
        ("softeningCells", C.c_float),
        ("timeScale", C.c_float),
        ("sortInterval", C.c_int),
        ("lengthScale", C.c_float),
        ("timeUnitScale", C.c_float),
        ("softBoundR", C.c_float),
        ("darkEnergy", C.c_float),
        ("contactEnabled", C.c_int),
        ("contactStiffness", C.c_float),
        ("contactDamping", C.c_float),
        ("pressureEnabled", C.c_int),
        ("coolingEnabled", C.c_int),
        ("starFormationEnabled", C.c_int),
        ("blackHoleEnabled", C.c_int),
        ("strongForceEnabled", C.c_int),
        ("emForceEnabled", C.c_int),
        ("weakForceEnabled", C.c_int),
        ("collapseEnabled", C.c_int),
        ("contactRadiusCells", C.c_float),
        ("pressureK", C.c_float),
        ("coolingRate", C.c_float),
    ]


class Modes(C.Structure):
    _fields_ = [
        ("keyMode", C.c_int),
        ("depositMode", C.c_int),
        ("integrator", C.c_int),
        ("pipeline", C.c_int),
        ("momentumZero", C.c_int),
        ("exactGrid", C.c_int),
        ("sortMerge", C.c_int),
    ]


class Timings(C.Structure):
    """sd_timings == SimTimings (Sim.h:844-852)."""

    _fields_ = [
        ("scatterMs", C.c_float),
        ("poissonMs", C.c_float),
        ("gatherMs", C.c_float),
        ("totalMs", C.c_float),
        ("dtUsed", C.c_float),
        ("maxSpeed", C.c_float),
    ]


_lib = None


def lib():
    """Load the HIP library; raises (never falls back) when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise StardustError(
                f"{LIB_PATH} is missing: build it with `python -m stardust_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.c_void_p
    sig = {
        "sd_config_default": (None, [C.POINTER(Config)]),
        "sd_modes_default": (None, [C.POINTER(Modes)]),
        "sd_device_available": (C.c_int, []),
        "sd_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
        "sd_device_driver": (C.c_int, [C.c_char_p, C.c_size_t]),
        "sd_device_free_bytes": (C.c_size_t, []),
        "sd_device_multiprocessors": (C.c_int, []),
        "sd_max_grid_size": (C.c_int, [C.c_int]),
        "sd_estimate_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
        "sd_max_particles_for": (C.c_int, [C.c_int, C.c_int, C.c_size_t]),
        "sd_device_bandwidth_gbs": (C.c_double, []),
        "sd_estimate_step_ms": (C.c_double, [C.c_int, C.c_int, C.c_int]),
        "sd_failed": (C.c_int, []),
        "sd_fail_message": (C.c_char_p, []),
        "sd_clear_failure": (None, []),
        "sd_create": (P, []),
        "sd_destroy": (None, [P]),
        "sd_set_modes": (C.c_int, [P, C.POINTER(Modes)]),
        "sd_get_modes": (C.c_int, [P, C.POINTER(Modes)]),
        "sd_init": (C.c_int, [P, C.POINTER(Config)]),
        "sd_reconfigure": (C.c_int, [P, C.POINTER(Config)]),
        "sd_reset": (C.c_int, [P]),
        "sd_step": (C.c_int, [P]),
        "sd_run": (C.c_int, [P, C.c_int]),
        "sd_synchronize": (C.c_int, [P]),
        "sd_get_config": (C.c_int, [P, C.POINTER(Config)]),
        "sd_timings_get": (C.c_int, [P, C.POINTER(Timings)]),
        "sd_sim_time": (C.c_double, [P]),
        "sd_step_count": (C.c_longlong, [P]),
        "sd_grid_size": (C.c_int, [P]),
        "sd_particle_count": (C.c_int, [P]),
        "sd_active_count": (C.c_int, [P]),
        "sd_measure_total_grid_mass": (C.c_double, [P]),
        "sd_measure_max_density": (C.c_double, [P]),
        "sd_measure_occupied_cells": (C.c_int, [P]),
        "sd_measure_centroid": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "sd_measure_kinetic_energy": (C.c_double, [P]),
        "sd_peak_cell_count": (C.c_int, [P]),
        "sd_save_state": (C.c_int, [P, C.c_char_p]),
        "sd_load_state": (C.c_int, [P, C.c_char_p]),
        "sd_set_particles": (C.c_int, [P, P, P, C.c_int]),
        "sd_get_particles": (C.c_int, [P, P, P, C.c_int]),
        "sd_set_fixed_dt": (C.c_int, [P, C.c_float]),
        "sd_particle_pos_device_ptr": (P, [P]),
        "sd_particle_vel_device_ptr": (P, [P]),
        "sd_get_grid": (C.c_int, [P, C.c_int, P, C.c_size_t]),
        "sd_compute_forces": (C.c_int, [P]),
        "sd_sort_now": (C.c_int, [P, P, P]),
        "sd_particle_accel": (C.c_int, [P, P, C.c_int]),
        "sd_direct_sum": (C.c_int, [P, P, C.c_int, C.c_double, C.c_double, P]),
        "sd_direct_sum_noalloc": (C.c_int, [P, P, C.c_int, C.c_double, C.c_double, P]),
        "sd_density_device_ptr": (C.c_void_p, [P]),
        "sd_field_device_ptr": (C.c_void_p, [P, C.c_int]),
        "sd_mlp_param_counts": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "sd_mlp_create": (P, [C.c_uint]),
        "sd_mlp_destroy": (None, [P]),
        "sd_mlp_set_params": (C.c_int, [P, P, P]),
        "sd_mlp_get_params": (C.c_int, [P, P, P]),
        "sd_mlp_forward": (C.c_int, [P, P, C.c_int, P]),
        "sd_mlp_grad": (C.c_int, [P, P, P, C.c_int, C.POINTER(C.c_double), P, P]),
        "sd_mlp_train": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, P]),
        "sd_features": (C.c_int, [P, P, C.c_int, P]),
        "sd_attach_mlp": (C.c_int, [P, P, C.c_float]),
        "sd_mlp_correction": (C.c_int, [P, P, C.c_int]),
        "sd_set_mlp_interval": (C.c_int, [P, C.c_int]),
        "sd_batch_create": (P, []),
        "sd_batch_destroy": (None, [P]),
        "sd_batch_add_scene": (C.c_int, [P, C.POINTER(Config)]),
        "sd_batch_init": (C.c_int, [P]),
        "sd_batch_scene_count": (C.c_int, [P]),
        "sd_batch_scene": (P, [P, C.c_int]),
        "sd_batch_total_particles": (C.c_longlong, [P]),
        "sd_batch_run": (C.c_int, [P, C.c_int, C.c_int]),
        "sd_batch_synchronize": (C.c_int, [P]),
        "sd_batch_last_run_ms": (C.c_float, [P]),
        "sd_batch_graph_nodes": (C.c_longlong, [P]),
        "sd_batch_snapshot_count": (C.c_int, [P]),
        "sd_batch_get_snapshot": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int]),
        "sd_kernel_timer_enable": (C.c_int, [P, C.c_char_p]),
        "sd_kernel_timer_every": (C.c_int, [P, C.c_int]),
        "sd_kernel_time_ms": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _check(rc, what):
    if rc != SD_OK:
        msg = lib().sd_fail_message()
        raise StardustError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def default_config(**kw) -> Config:
    """SimConfig defaults with the plain-PM switches set (Empty preset, no
    off-path physics), then overridden by keyword arguments."""
    c = Config()
    lib().sd_config_default(C.byref(c))
    c.preset = PRESET_EMPTY
    c.pressureEnabled = 0
    c.coolingEnabled = 0
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def default_modes(**kw) -> Modes:
    m = Modes()
    lib().sd_modes_default(C.byref(m))
    for k, v in kw.items():
        if not hasattr(m, k):
            raise KeyError(k)
        setattr(m, k, v)
    return m


def device_available() -> bool:
    return bool(lib().sd_device_available())


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Sim:
    """Python mirror of the reference's ``class Sim`` (Sim.h:854-1135)."""

    def __init__(self, config: Config | None = None, modes: Modes | None = None):
        self._h = lib().sd_create()
        if not self._h:
            raise StardustError("sd_create failed")
        if modes is not None:
            _check(lib().sd_set_modes(self._h, C.byref(modes)), "sd_set_modes")
        if config is not None:
            self.init(config)

    @classmethod
    def _view(cls, handle, owner):
        """A Sim over a handle owned elsewhere (a batch scene): never destroyed here."""
        self = cls.__new__(cls)
        self._h = handle
        self._owner = owner
        return self

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                lib().sd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # lifecycle (Sim.h:891-895)
    def init(self, config: Config):
        _check(lib().sd_init(self._h, C.byref(config)), "sd_init")

    def reconfigure(self, config: Config):
        _check(lib().sd_reconfigure(self._h, C.byref(config)), "sd_reconfigure")

    def reset(self):
        _check(lib().sd_reset(self._h), "sd_reset")

    def step(self):
        _check(lib().sd_step(self._h), "sd_step")

    def run(self, steps: int):
        _check(lib().sd_run(self._h, int(steps)), "sd_run")

    def synchronize(self):
        _check(lib().sd_synchronize(self._h), "sd_synchronize")

    def set_modes(self, modes: Modes):
        _check(lib().sd_set_modes(self._h, C.byref(modes)), "sd_set_modes")

    def modes(self) -> Modes:
        m = Modes()
        _check(lib().sd_get_modes(self._h, C.byref(m)), "sd_get_modes")
        return m

    # reads
    def config(self) -> Config:
        c = Config()
        _check(lib().sd_get_config(self._h, C.byref(c)), "sd_get_config")
        return c

    def timings(self) -> Timings:
        t = Timings()
        _check(lib().sd_timings_get(self._h, C.byref(t)), "sd_timings_get")
        return t

    def sim_time(self) -> float:
        return float(lib().sd_sim_time(self._h))

    def step_count(self) -> int:
        return int(lib().sd_step_count(self._h))

    def grid_size(self) -> int:
        return int(lib().sd_grid_size(self._h))

    def particle_count(self) -> int:
        return int(lib().sd_particle_count(self._h))

    def active_count(self) -> int:
        return int(lib().sd_active_count(self._h))

    # measures (Sim.h:1000-1029)
    def measure_total_grid_mass(self) -> float:
        return float(lib().sd_measure_total_grid_mass(self._h))

    def measure_max_density(self) -> float:
        return float(lib().sd_measure_max_density(self._h))

    def measure_occupied_cells(self) -> int:
        return int(lib().sd_measure_occupied_cells(self._h))

    def measure_centroid(self):
        cx, cy = C.c_double(), C.c_double()
        _check(lib().sd_measure_centroid(self._h, C.byref(cx), C.byref(cy)), "measure_centroid")
        return cx.value, cy.value

    def measure_kinetic_energy(self) -> float:
        return float(lib().sd_measure_kinetic_energy(self._h))

    def peak_cell_count(self) -> int:
        return int(lib().sd_peak_cell_count(self._h))

    # state files
    def save_state(self, path: str):
        _check(lib().sd_save_state(self._h, path.encode()), "sd_save_state")

    def load_state(self, path: str):
        _check(lib().sd_load_state(self._h, path.encode()), "sd_load_state")

    # particles
    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        assert pos.shape == vel.shape and pos.ndim == 2 and pos.shape[1] == 4
        _check(lib().sd_set_particles(self._h, _ptr(pos), _ptr(vel), pos.shape[0]),
               "sd_set_particles")

    def get_particles(self, n: int | None = None):
        n = self.particle_count() if n is None else n
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        _check(lib().sd_get_particles(self._h, _ptr(pos), _ptr(vel), n), "sd_get_particles")
        return pos, vel

    def set_fixed_dt(self, dt: float):
        _check(lib().sd_set_fixed_dt(self._h, float(dt)), "sd_set_fixed_dt")

    def get_grid(self, which: int):
        G = self.grid_size()
        if which == GRID_DENSITY_FIXED:
            out = np.empty((G, G, G), np.int64)
        elif which == GRID_ACCEL:
            out = np.empty((G, G, G, 4), np.float32)
        else:
            out = np.empty((G, G, G), np.float32)
        _check(lib().sd_get_grid(self._h, which, _ptr(out), out.nbytes), "sd_get_grid")
        return out

    def density_device_ptr(self) -> int:
        """Sim::densityDevicePtr: device address of float[G^3] particles per cell."""
        p = lib().sd_density_device_ptr(self._h)
        if not p:
            raise StardustError("sd_density_device_ptr: " + fail_message())
        return p

    def field_device_ptr(self, field: int = 0) -> int:
        """Sim::fieldDevicePtr: device address of float[G^2] (density summed along z)."""
        p = lib().sd_field_device_ptr(self._h, field)
        if not p:
            raise StardustError("sd_field_device_ptr: " + fail_message())
        return p

    def compute_forces(self):
        _check(lib().sd_compute_forces(self._h), "sd_compute_forces")

    def sort_now(self):
        n = self.particle_count()
        keys = np.empty(n, np.uint32)
        order = np.empty(n, np.int32)
        _check(lib().sd_sort_now(self._h, _ptr(keys), _ptr(order)), "sd_sort_now")
        return keys, order

    def particle_accel(self, n: int | None = None):
        n = self.particle_count() if n is None else n
        out = np.empty((n, 4), np.float32)
        _check(lib().sd_particle_accel(self._h, _ptr(out), n), "sd_particle_accel")
        return out

    def direct_sum(self, targets, gm: float, eps: float):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        out = np.empty((t.shape[0], 3), np.float64)
        _check(lib().sd_direct_sum(self._h, _ptr(t), t.shape[0], gm, eps, _ptr(out)),
               "sd_direct_sum")
        return out

    def features(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        out = np.empty((t.shape[0], 30), np.float32)
        _check(lib().sd_features(self._h, _ptr(t), t.shape[0], _ptr(out)), "sd_features")
        return out

    def attach_mlp(self, mlp: "Mlp | None", out_scale: float = 1.0):
        _check(lib().sd_attach_mlp(self._h, mlp._h if mlp is not None else None, float(out_scale)),
               "sd_attach_mlp")
        self._mlp = mlp  # keep it alive while attached

    def set_mlp_interval(self, steps: int):
        _check(lib().sd_set_mlp_interval(self._h, int(steps)), "sd_set_mlp_interval")

    def mlp_correction(self, n: int | None = None):
        n = self.particle_count() if n is None else n
        out = np.empty((n, 4), np.float32)
        _check(lib().sd_mlp_correction(self._h, _ptr(out), n), "sd_mlp_correction")
        return out

    def kernel_timer(self, name: str | None, every: int = 1):
        """Time launches of one named kernel with HIP events (one launch in `every`)."""
        _check(lib().sd_kernel_timer_enable(self._h, (name or "").encode()), "kernel_timer")
        _check(lib().sd_kernel_timer_every(self._h, int(every)), "kernel_timer_every")

    def kernel_time_ms(self):
        ms, n = C.c_double(), C.c_longlong()
        _check(lib().sd_kernel_time_ms(self._h, C.byref(ms), C.byref(n)), "kernel_time_ms")
        return ms.value, n.value


class Mlp:
    """The learned correction network (sd_mlp): 30 -> 128 -> 128 -> 128 -> 3."""

    def __init__(self, seed: int = 1):
        self._h = lib().sd_mlp_create(int(seed))
        if not self._h:
            raise StardustError("sd_mlp_create failed: " + fail_message())
        nw, nb = C.c_int(), C.c_int()
        lib().sd_mlp_param_counts(C.byref(nw), C.byref(nb))
        self.nW, self.nB = nw.value, nb.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        W = np.empty(self.nW, np.float32)
        b = np.empty(self.nB, np.float32)
        _check(lib().sd_mlp_get_params(self._h, _ptr(W), _ptr(b)), "sd_mlp_get_params")
        return W, b

    def set_params(self, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert W.size == self.nW and b.size == self.nB
        _check(lib().sd_mlp_set_params(self._h, _ptr(W), _ptr(b)), "sd_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty((X.shape[0], 3), np.float32)
        _check(lib().sd_mlp_forward(self._h, _ptr(X), X.shape[0], _ptr(Y)), "sd_mlp_forward")
        return Y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        gW = np.empty(self.nW, np.float32)
        gb = np.empty(self.nB, np.float32)
        loss = C.c_double()
        _check(lib().sd_mlp_grad(self._h, _ptr(X), _ptr(T), X.shape[0], C.byref(loss), _ptr(gW),
                                 _ptr(gb)), "sd_mlp_grad")
        return loss.value, gW, gb

    def train(self, X, T, epochs, batch=8192, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        losses = np.zeros(max(epochs, 1), np.float32)
        _check(lib().sd_mlp_train(self._h, _ptr(X), _ptr(T), X.shape[0], int(epochs), int(batch),
                                  float(lr), int(seed), _ptr(losses)), "sd_mlp_train")
        return losses[:epochs]


def failed() -> bool:
    return bool(lib().sd_failed())


def fail_message() -> str:
    m = lib().sd_fail_message()
    return m.decode() if m else ""


def clear_failure():
    lib().sd_clear_failure()


class Batch:
    """Many independent scenes stepped together (one launch per kernel for all
    scenes, a whole run recorded into one HIP graph). ``scene(i)`` is a
    :class:`Sim` view of scene i for particle I/O and measures."""

    def __init__(self, configs):
        self._h = lib().sd_batch_create()
        if not self._h:
            raise StardustError("sd_batch_create failed")
        for c in configs:
            rc = lib().sd_batch_add_scene(self._h, C.byref(c))
            if rc < 0:
                _check(rc, "sd_batch_add_scene")
        _check(lib().sd_batch_init(self._h), "sd_batch_init")
        self.scenes = [Sim._view(lib().sd_batch_scene(self._h, i), self) for i in range(len(configs))]

    def close(self):
        if getattr(self, "_h", None):
            for s in self.scenes:
                s._h = None
            lib().sd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scene(self, i) -> Sim:
        return self.scenes[i]

    @property
    def total_particles(self) -> int:
        return int(lib().sd_batch_total_particles(self._h))

    def run(self, steps: int, snapshot_every: int = 0):
        _check(lib().sd_batch_run(self._h, int(steps), int(snapshot_every)), "sd_batch_run")

    def synchronize(self):
        _check(lib().sd_batch_synchronize(self._h), "sd_batch_synchronize")

    def last_run_ms(self) -> float:
        return float(lib().sd_batch_last_run_ms(self._h))

    def graph_nodes(self) -> int:
        return int(lib().sd_batch_graph_nodes(self._h))

    def snapshot_count(self) -> int:
        return int(lib().sd_batch_snapshot_count(self._h))

    def snapshot(self, snap: int, scene: int):
        n = self.scenes[scene].particle_count()
        out = np.empty((n, 4), np.float32)
        _check(lib().sd_batch_get_snapshot(self._h, snap, scene, out.ctypes.data, n), "sd_batch_get_snapshot")
        return out
