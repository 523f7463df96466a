"""ctypes front-end of the CPU oracle (``libsd_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker. The product path
(``stardust_amd``) never imports it.

Every wrapper returns numpy arrays; particle arrays are ``float32[n, 4]``
(x, y, z, w) in box units with ``x < 0`` marking an inactive slot.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsd_oracle.so")
_lib = None

FIX_SHIFT = 24
FIX_ONE = 1 << FIX_SHIFT
KEY_ROWMAJOR = 0
KEY_MORTON = 1
KEY_MORTON_NODE = 2


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("G", C.c_int),
        ("periodic", C.c_int),
        ("gravity", C.c_float),
        ("softeningCells", C.c_float),
        ("timeScale", C.c_float),
        ("sortInterval", C.c_int),
        ("keyMode", C.c_int),
        ("fixedDeposit", C.c_int),
        ("fixedDt", C.c_float),
        ("momentumZero", C.c_int),
        ("contact", C.c_int),
        ("contactRadiusCells", C.c_float),
        ("contactStiffness", C.c_float),
        ("contactDamping", C.c_float),
        ("c2", C.c_float),
        ("softBoundR", C.c_float),
        ("darkEnergy", C.c_float),
        ("allocN", C.c_int),
        ("fp32Mesh", C.c_int),
        ("pressureEnabled", C.c_int),
        ("coolingEnabled", C.c_int),
        ("pressureK", C.c_float),
        ("coolingRate", C.c_float),
    ]


class State(C.Structure):
    _fields_ = [
        ("stepCount", C.c_longlong),
        ("simTime", C.c_double),
        ("lastDt", C.c_float),
        ("lastVmax", C.c_float),
    ]


def _declare(L):
    L.sdo_morton3.restype = C.c_uint32
    L.sdo_morton3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.sdo_cell_keys.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _u32p, _i32p]
    L.sdo_stable_sort_pairs.argtypes = [_u32p, _i32p, C.c_int]
    L.sdo_reorder4.argtypes = [_fp, _i32p, C.c_int, _fp]
    L.sdo_deposit_float.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, _fp]
    L.sdo_deposit_fixed.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _i64p]
    L.sdo_cic_weights_fixed.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _i32p, _i64p]
    L.sdo_fft1d.argtypes = [_dp, _dp, C.c_int, C.c_int]
    L.sdo_fft3d.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.sdo_green_real.argtypes = [C.c_int, C.c_float, _dp]
    L.sdo_poisson_iso.argtypes = [_dp, C.c_int, C.c_float, _dp]
    L.sdo_poisson_periodic.argtypes = [_dp, C.c_int, C.c_float, _dp]
    L.sdo_poisson_iso_f32.argtypes = [_fp, C.c_int, C.c_float, _fp]
    L.sdo_poisson_periodic_f32.argtypes = [_fp, C.c_int, C.c_float, _fp]
    L.sdo_grid_accel.argtypes = [_dp, C.c_int, C.c_float, C.c_int, _fp]
    L.sdo_sample_acc.argtypes = [_fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _fp]
    L.sdo_sample_acc_all.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, _fp]
    L.sdo_max_speed.restype = C.c_float
    L.sdo_max_speed.argtypes = [_fp, _fp, C.c_int]
    L.sdo_integrate.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                C.c_float, C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    L.sdo_cool_cells.argtypes = [_fp, _fp, _u32p, C.c_int, C.c_int, C.c_float, C.c_float, _fp]
    L.sdo_pressure.argtypes = [_fp, C.c_int, C.c_int, C.c_float, C.c_float, _fp]
    L.sdo_momentum.argtypes = [_fp, _fp, C.c_int, _dp]
    L.sdo_subtract_mean.argtypes = [_fp, _fp, C.c_int, _dp]
    L.sdo_cell_range.argtypes = [_u32p, C.c_int, C.c_int, _i32p, _i32p]
    L.sdo_contact.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p,
                              C.c_float, C.c_float, C.c_float, _fp]
    L.sdo_dt_policy.restype = C.c_float
    L.sdo_dt_policy.argtypes = [C.POINTER(Params), C.c_float]
    L.sdo_compute_accel.argtypes = [C.POINTER(Params), _fp, C.c_int, _fp, C.c_void_p, C.c_void_p]
    L.sdo_step.restype = C.c_int
    L.sdo_step.argtypes = [C.POINTER(Params), C.POINTER(State), _fp, _fp, C.c_int, _fp, _fp, C.c_void_p]
    L.sdo_direct_sum.argtypes = [_fp, C.c_int, _i32p, C.c_int, C.c_double, C.c_double, _dp]
    L.sdo_rel_rms.restype = C.c_double
    L.sdo_rel_rms.argtypes = [_dp, _dp, C.c_int, C.POINTER(C.c_double)]
    L.sdo_mlp_forward.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, _dp]
    L.sdo_mlp_loss_grad.restype = C.c_double
    L.sdo_mlp_loss_grad.argtypes = [_dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.sdo_mlp_wcount.restype = C.c_size_t
    L.sdo_mlp_wcount.argtypes = [C.c_int]
    L.sdo_mlp_bcount.restype = C.c_size_t
    L.sdo_mlp_bcount.argtypes = [C.c_int]


def _p4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def morton3(x, y, z):
    return int(lib().sdo_morton3(x, y, z))


def cell_keys(pos, G, mode=KEY_ROWMAJOR):
    pos = _p4(pos)
    n = pos.shape[0]
    keys = np.empty(n, np.uint32)
    idx = np.empty(n, np.int32)
    lib().sdo_cell_keys(pos, n, G, mode, keys, idx)
    return keys, idx


def stable_sort_pairs(keys, vals):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    v = np.ascontiguousarray(vals, dtype=np.int32).copy()
    lib().sdo_stable_sort_pairs(k, v, k.shape[0])
    return k, v


def reorder4(src, order):
    src = _p4(src)
    order = np.ascontiguousarray(order, dtype=np.int32)
    dst = np.empty_like(src)
    lib().sdo_reorder4(src, order, order.shape[0], dst)
    return dst


def deposit_float(pos, G, S=None, periodic=False):
    pos = _p4(pos)
    S = G if S is None else S
    rho = np.empty(S * S * S, np.float32)
    lib().sdo_deposit_float(pos, pos.shape[0], G, S, int(periodic), rho)
    return rho.reshape(S, S, S)


def deposit_fixed(pos, G, periodic=False):
    pos = _p4(pos)
    rho = np.empty(G * G * G, np.int64)
    lib().sdo_deposit_fixed(pos, pos.shape[0], G, int(periodic), rho)
    return rho.reshape(G, G, G)


def cic_weights_fixed(p, G):
    b = np.empty(3, np.int32)
    w = np.empty(8, np.int64)
    lib().sdo_cic_weights_fixed(float(p[0]), float(p[1]), float(p[2]), G, b, w)
    return b, w


def fft1d(x, inverse=False):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib().sdo_fft1d(re, im, x.shape[0], int(inverse))
    return re + 1j * im


def fft3d(x, inverse=False):
    """x laid out [z][y][x]."""
    x = np.asarray(x, dtype=np.complex128)
    nz, ny, nx = x.shape
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib().sdo_fft3d(re.reshape(-1), im.reshape(-1), nx, ny, nz, int(inverse))
    return re + 1j * im


def green_real(S, soft):
    g = np.empty(S * S * S, np.float64)
    lib().sdo_green_real(S, soft, g)
    return g.reshape(S, S, S)


def poisson_iso(rho, softCells):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    G = rho.shape[0]
    pot = np.empty(G * G * G, np.float64)
    lib().sdo_poisson_iso(rho.reshape(-1), G, softCells, pot)
    return pot.reshape(G, G, G)


def poisson_periodic(rho, softCells):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    G = rho.shape[0]
    pot = np.empty(G * G * G, np.float64)
    lib().sdo_poisson_periodic(rho.reshape(-1), G, softCells, pot)
    return pot.reshape(G, G, G)


def poisson_iso_f32(rho, softCells):
    """Single-precision isolated solve (the CPU baseline's mesh)."""
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    G = rho.shape[0]
    pot = np.empty(G * G * G, np.float32)
    lib().sdo_poisson_iso_f32(rho.reshape(-1), G, softCells, pot)
    return pot.reshape(G, G, G)


def poisson_periodic_f32(rho, softCells):
    """Single-precision periodic solve (the CPU baseline's mesh)."""
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    G = rho.shape[0]
    pot = np.empty(G * G * G, np.float32)
    lib().sdo_poisson_periodic_f32(rho.reshape(-1), G, softCells, pot)
    return pot.reshape(G, G, G)


def grid_accel(pot, potScale, periodic=False):
    pot = np.ascontiguousarray(pot, dtype=np.float64)
    G = pot.shape[0]
    acc = np.empty(G * G * G * 4, np.float32)
    lib().sdo_grid_accel(pot.reshape(-1), G, potScale, int(periodic), acc)
    return acc.reshape(G, G, G, 4)


def sample_acc(acc4, p, G, periodic=False):
    acc4 = np.ascontiguousarray(acc4, dtype=np.float32).reshape(-1)
    out = np.empty(3, np.float32)
    lib().sdo_sample_acc(acc4, float(p[0]), float(p[1]), float(p[2]), G, int(periodic), out)
    return out


def particle_accel(params, pos):
    """Mesh acceleration at each particle (compute_accel, then the CIC gather);
    zero for inactive slots. Returns float32 [n, 4]."""
    pos = _p4(pos)
    acc = compute_accel(params, pos)
    flat = np.ascontiguousarray(acc, dtype=np.float32).reshape(-1)
    out = np.empty((pos.shape[0], 4), np.float32)
    lib().sdo_sample_acc_all(flat, pos, pos.shape[0], params.G, int(params.periodic), out)
    return out


def max_speed(pos, vel):
    pos, vel = _p4(pos), _p4(vel)
    return float(lib().sdo_max_speed(pos, vel, pos.shape[0]))


def integrate(acc4, pos, vel, G, kick, drift, periodic=False, c2=1e4, acc_contact=None,
              softBoundR=0.0, darkEnergy=0.0, mean_sub=None, acc_press=None):
    """Returns updated (pos, vel) copies."""
    acc4 = np.ascontiguousarray(acc4, dtype=np.float32).reshape(-1)
    pos, vel = _p4(pos).copy(), _p4(vel).copy()
    ac = None
    if acc_contact is not None:
        ac = _p4(acc_contact)
    ms = None
    if mean_sub is not None:
        ms = np.ascontiguousarray(mean_sub, dtype=np.float32)
    ap = None
    if acc_press is not None:
        ap = np.ascontiguousarray(acc_press, dtype=np.float32).reshape(-1)
    lib().sdo_integrate(acc4, pos, vel, pos.shape[0], G, kick, drift, int(periodic), c2,
                        ac.ctypes.data if ac is not None else None, softBoundR, darkEnergy,
                        ms.ctypes.data if ms is not None else None,
                        ap.ctypes.data if ap is not None else None)
    return pos, vel


def cool_cells(pos, vel, keys, G, rate, dt):
    """Cooling pass over particles sorted by cell; returns (vel, disp[4, G, G, G])."""
    pos, vel = _p4(pos), _p4(vel).copy()
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    disp = np.zeros(4 * G ** 3, np.float32)
    lib().sdo_cool_cells(pos, vel, k, pos.shape[0], G, rate, dt, disp)
    return vel, disp.reshape(4, G, G, G)


def pressure(disp, G, periodic, k, max_acc=5.0):
    """kPressure of the dispersion grids disp[4, G, G, G] -> float32 [G, G, G, 4]."""
    d = np.ascontiguousarray(disp, dtype=np.float32).reshape(-1)
    out = np.empty(4 * G ** 3, np.float32)
    lib().sdo_pressure(d, G, int(periodic), k, max_acc, out)
    return out.reshape(G, G, G, 4)


def momentum(pos, vel):
    out = np.empty(4, np.float64)
    pos, vel = _p4(pos), _p4(vel)
    lib().sdo_momentum(pos, vel, pos.shape[0], out)
    return out


def subtract_mean(pos, vel, acc4):
    pos, vel = _p4(pos), _p4(vel).copy()
    lib().sdo_subtract_mean(pos, vel, pos.shape[0], np.ascontiguousarray(acc4, np.float64))
    return vel


def cell_range(sorted_keys, ncells):
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    s = np.empty(ncells, np.int32)
    e = np.empty(ncells, np.int32)
    lib().sdo_cell_range(k, k.shape[0], ncells, s, e)
    return s, e


def contact(pos, vel, G, cell_start, cell_end, radius, stiffness, damping, periodic=False,
            key_mode=KEY_ROWMAJOR):
    pos, vel = _p4(pos), _p4(vel)
    out = np.empty_like(pos)
    lib().sdo_contact(pos, vel, pos.shape[0], G, int(periodic), key_mode,
                      np.ascontiguousarray(cell_start, np.int32),
                      np.ascontiguousarray(cell_end, np.int32), radius, stiffness, damping, out)
    return out


def make_params(G, periodic=False, gravity=1.0, softeningCells=3.0, timeScale=1.0,
                sortInterval=40, keyMode=KEY_ROWMAJOR, fixedDeposit=1, fixedDt=0.0,
                momentumZero=1, contact=0, contactRadiusCells=0.5, contactStiffness=1e6,
                contactDamping=0.35, c2=1e4, softBoundR=0.0, darkEnergy=0.0, allocN=1, fp32Mesh=0,
                pressure=0, cooling=0, pressureK=0.45, coolingRate=0.08):
    return Params(G, int(periodic), gravity, softeningCells, timeScale, sortInterval, keyMode,
                  fixedDeposit, fixedDt, momentumZero, contact, contactRadiusCells,
                  contactStiffness, contactDamping, c2, softBoundR, darkEnergy, allocN, int(fp32Mesh),
                  int(pressure), int(cooling), pressureK, coolingRate)


def dt_policy(params, vmax):
    return float(lib().sdo_dt_policy(C.byref(params), vmax))


def compute_accel(params, pos, want_rho=False, want_pot=False):
    pos = _p4(pos)
    G = params.G
    acc = np.empty(G * G * G * 4, np.float32)
    rho = np.empty(G ** 3, np.int64) if want_rho else None
    pot = np.empty(G ** 3, np.float64) if want_pot else None
    lib().sdo_compute_accel(C.byref(params), pos, pos.shape[0], acc,
                            rho.ctypes.data if rho is not None else None,
                            pot.ctypes.data if pot is not None else None)
    out = [acc.reshape(G, G, G, 4)]
    if want_rho:
        out.append(rho.reshape(G, G, G))
    if want_pot:
        out.append(pot.reshape(G, G, G))
    return out[0] if len(out) == 1 else tuple(out)


class Stepper:
    """Runs the oracle's reference-order ``Sim::step`` on host copies."""

    def __init__(self, params, pos, vel):
        self.params = params
        self.state = State(0, 0.0, 0.0, 0.0)
        self.pos = _p4(pos).copy()
        self.vel = _p4(vel).copy()
        self._tp = np.empty_like(self.pos)
        self._tv = np.empty_like(self.vel)
        # gas dispersion grids carried from one cooling pass to the next steps
        self.gas = None
        if params.pressureEnabled or params.coolingEnabled:
            self.gas = np.zeros(4 * params.G ** 3, np.float32)

    def step(self, k=1):
        for _ in range(k):
            lib().sdo_step(C.byref(self.params), C.byref(self.state), self.pos, self.vel,
                           self.pos.shape[0], self._tp, self._tv,
                           self.gas.ctypes.data if self.gas is not None else None)


def direct_sum(pos, targets, gm, eps):
    pos = _p4(pos)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    out = np.empty(t.shape[0] * 3, np.float64)
    lib().sdo_direct_sum(pos, pos.shape[0], t, t.shape[0], gm, eps, out)
    return out.reshape(-1, 3)


def rel_rms(approx, truth):
    a = np.ascontiguousarray(approx, np.float64).reshape(-1)
    t = np.ascontiguousarray(truth, np.float64).reshape(-1)
    s = C.c_double(0.0)
    r = lib().sdo_rel_rms(a, t, a.shape[0] // 3, C.byref(s))
    return float(r), float(s.value)


def mlp_counts(H):
    return int(lib().sdo_mlp_wcount(H)), int(lib().sdo_mlp_bcount(H))


def mlp_forward(W, b, H, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(x.shape[0] * 3, np.float64)
    lib().sdo_mlp_forward(np.ascontiguousarray(W, np.float64), np.ascontiguousarray(b, np.float64),
                          H, x.reshape(-1), x.shape[0], y)
    return y.reshape(-1, 3)


def mlp_loss_grad(W, b, H, x, target):
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(target, np.float64)
    nw, nb = mlp_counts(H)
    gW = np.empty(nw, np.float64)
    gb = np.empty(nb, np.float64)
    loss = lib().sdo_mlp_loss_grad(np.ascontiguousarray(W, np.float64),
                                   np.ascontiguousarray(b, np.float64), H, x.reshape(-1),
                                   t.reshape(-1), x.shape[0], gW, gb)
    return float(loss), gW, gb
