"""Batched ensemble of independent traps (am_ensemble_* in the C ABI).

Each trap is one instance of the reference's Simulate loop (src/antimatter-kernel.cc:153-197)
with its own field, dt, cooling and annihilation settings; all traps advance together in one
store.  ``make_config5`` draws the 512-trap ensemble of the batched workload.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import Context, Store, _check, _p, lib, make_trap
from .driver import derive_dt, initial_conditions

_d = ctypes.c_double
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p

TRAP_SEED_STEP = 0x9E3779B97F4A7C15


class TrapDesc(ctypes.Structure):
    _fields_ = [("begin", _i64), ("end", _i64), ("b0", _d), ("trap_voltage", _d), ("d", _d), ("dt", _d),
                ("tau_scale", _d), ("bath_temp", _d), ("annihilation", _i32), ("contact_radius", _d),
                ("probability", _d)]


SIGNATURES = {
    "am_ensemble_create": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp]),
    "am_ensemble_destroy": (ctypes.c_int, [_vp]),
    "am_ensemble_set_chunk": (ctypes.c_int, [_vp, _i32]),
    "am_ensemble_autotune": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "am_ensemble_run": (ctypes.c_int, [_vp, _i32, ctypes.c_uint64, _i32]),
    "am_ensemble_stats": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
}


def _bind():
    L = lib()
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def trap_seed(seed: int, t: int) -> int:
    """Key of trap t's random draws (seed + t * 0x9E3779B97F4A7C15 mod 2^64)."""
    return (seed + t * TRAP_SEED_STEP) % (1 << 64)


class Ensemble:
    def __init__(self, store: Store, traps: list, contact_capacity: int = 0):
        L = _bind()
        arr = (TrapDesc * len(traps))(*traps)
        h = ctypes.c_void_p()
        _check(L.am_ensemble_create(store.h, len(traps), ctypes.cast(arr, _vp), int(contact_capacity),
                                    ctypes.byref(h)))
        self.h = h
        self.store = store
        self.n_traps = len(traps)
        store._children.add(self)

    def close(self):
        if self.h:
            if self.store.h:
                lib().am_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_chunk(self, chunk: int):
        _check(lib().am_ensemble_set_chunk(self.h, int(chunk)))

    def autotune(self, candidates=None):
        c = None if candidates is None else np.asarray(candidates, dtype=np.int32)
        k = 6 if c is None else len(c)
        times = np.zeros(k)
        chosen = ctypes.c_int32()
        _check(lib().am_ensemble_autotune(self.h, _p(c), 0 if c is None else len(c), ctypes.byref(chosen),
                                          _p(times)))
        return chosen.value, times

    def run(self, steps: int, seed: int = 42, use_graph: bool = True):
        _check(lib().am_ensemble_run(self.h, int(steps), int(seed), int(use_graph)))

    def stats(self):
        sd, an, it = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        ch = ctypes.c_int32()
        pairs = ctypes.c_double()
        _check(lib().am_ensemble_stats(self.h, ctypes.byref(sd), ctypes.byref(an), ctypes.byref(ch), ctypes.byref(it),
                                       ctypes.byref(pairs)))
        return dict(steps_done=sd.value, annihilated=an.value, chunk=ch.value, n_items=it.value,
                    pair_interactions_per_step=pairs.value)


SPECIES4 = ("positron", "electron", "antiproton", "proton")


def make_config5(n_traps: int = 512, seed: int = 42, contact_radius: float = 5e-6, probability: float = 1.0,
                 bath_temp: float = 4.0, tau_scale: float = 1.0, n_min: int = 1024, n_max: int = 32768,
                 annihilation: bool = True, radii_range=(1e-4, 2e-3)):
    """The batched workload: per trap, drawn from one seed, N log-uniform in [n_min, n_max], species
    fractions Dirichlet(1,1,1,1) over {e+, e-, pbar, p}, T log-uniform in [4, 1e4] K (all species),
    B0 uniform in [1, 5] T, U0 uniform in [5, 100] V, cloud radii uniform in radii_range (the workload's
    [1e-4, 2e-3] m) per axis, z0 = rho0 = 0.01 m; positions and velocities from the reference
    generators with seed + t.  Returns (pos, vel, species, trap descriptors, per-trap parameter dicts)."""
    rng = np.random.default_rng(seed)
    pos_l, vel_l, sp_l, traps, params = [], [], [], [], []
    begin = 0
    for t in range(n_traps):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        frac = rng.dirichlet(np.ones(4))
        counts = np.floor(frac * n).astype(int)
        counts[np.argmax(frac)] += n - counts.sum()
        T = float(math.exp(rng.uniform(math.log(4.0), math.log(1e4))))
        b0 = float(rng.uniform(1.0, 5.0))
        u0 = float(rng.uniform(5.0, 100.0))
        radii = tuple(float(r) for r in rng.uniform(radii_range[0], radii_range[1], 3))
        pops = {name: int(c) for name, c in zip(SPECIES4, counts) if c > 0}
        temps = {name: T for name in pops}
        pos, vel, sp = initial_conditions(pops, temps, radii, seed + t)
        field = make_trap(b0, u0, 0.01, 0.01)
        dt = derive_dt(pops, b0)
        end = begin + len(sp)
        traps.append(TrapDesc(begin, end, field.b0, field.trap_voltage, field.d, dt, tau_scale, bath_temp,
                              int(annihilation), contact_radius, probability))
        params.append(dict(populations=pops, temperature=T, b0=b0, u0=u0, radii=radii, dt=dt, d=field.d))
        pos_l.append(pos)
        vel_l.append(vel)
        sp_l.append(sp)
        begin = end
    return np.concatenate(pos_l), np.concatenate(vel_l), np.concatenate(sp_l), traps, params
