"""params.json and the native driver loop (include/antimatter_amd/am_params.h).

``Params`` mirrors the C struct ``am_params``: the fields of the reference's PhysicsParams,
InteractionParams and StochasticParams (src/params/*.h).  ``load_params`` / ``parse_params`` read a
file or JSON text with the reference's keys and conversions, ``validate`` runs the reference's three
Validate() checks, and ``simulate`` runs the driver loop of src/antimatter-kernel.cc:42-217 in the
library (C++ host code over the HIP kernels), returning the energy rows it logs.
"""
from __future__ import annotations

import ctypes
import json

import numpy as np

from . import FP64, _check, lib

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p
METHOD_LEN = 32


class Params(ctypes.Structure):
    _fields_ = [("seed", _i32), ("b0", _d), ("trap_voltage", _d), ("z0", _d), ("rho0", _d), ("dt", _d),
                ("num_steps", _u64), ("bath_temp", _d), ("cooling_scale", _d), ("populations", _u64 * 6),
                ("population_listed", _i32 * 6), ("temperatures", _d * 6), ("temperature_listed", _i32 * 6),
                ("initial_cloud_radii", _d * 3), ("coulomb_enabled", _i32),
                ("coulomb_method", ctypes.c_char * METHOD_LEN), ("energy_method", ctypes.c_char * METHOD_LEN),
                ("cutoff_radius", _d), ("barnes_hut_theta", _d), ("fast_multipole_order", _d),
                ("fast_multipole_theta", _d), ("annihilation_enabled", _i32), ("contact_radius", _d),
                ("annihilation_probability", _d), ("loaded", _i32)]


class RunOptions(ctypes.Structure):
    _fields_ = [("device", _i32), ("precision", _i32), ("rng_seed", _u64), ("print", _i32),
                ("print_final_positions", _i32), ("export_interval", _i32), ("export_dir", ctypes.c_char_p)]


class EnergyRow(ctypes.Structure):
    _fields_ = [("step", _i64), ("kinetic", _d), ("trap_potential", _d), ("coulomb_potential", _d), ("total", _d),
                ("num_active_particles", _i64), ("fractional_drift", _d), ("absolute_drift", _d)]


SIGNATURES = {
    "am_params_defaults": (ctypes.c_int, [_vp]),
    "am_params_load": (ctypes.c_int, [ctypes.c_char_p, _vp]),
    "am_params_parse": (ctypes.c_int, [ctypes.c_char_p, _vp]),
    "am_params_validate": (ctypes.c_int, [_vp]),
    "am_params_method": (ctypes.c_int, [ctypes.c_char_p, _vp]),
    "am_simulate": (ctypes.c_int, [_vp, _vp, ctypes.c_char_p, _vp, _i64, _vp]),
    "am_export_vtu": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "am_export_pvd": (ctypes.c_int, [ctypes.c_char_p, _vp, _vp, _i64]),
}


def _bind():
    L = lib()
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def defaults() -> Params:
    p = Params()
    _check(_bind().am_params_defaults(ctypes.byref(p)))
    return p


def load_params(path: str) -> Params:
    """LoadParams of the three groups on a file (a missing file keeps the defaults, loaded = 0)."""
    p = Params()
    _check(_bind().am_params_load(path.encode(), ctypes.byref(p)))
    return p


def parse_params(config) -> Params:
    """LoadParams on JSON text, or on a dict written out as JSON."""
    text = config if isinstance(config, str) else json.dumps(config)
    p = Params()
    _check(_bind().am_params_parse(text.encode(), ctypes.byref(p)))
    return p


def validate(p: Params) -> None:
    """The reference's PhysicsParams/InteractionParams/StochasticParams::Validate (raises AmError)."""
    _check(_bind().am_params_validate(ctypes.byref(p)))


def method_code(name: str) -> int:
    m = ctypes.c_int32()
    _check(_bind().am_params_method(name.encode(), ctypes.byref(m)))
    return m.value


def simulate(config, rng_seed: int = 42, precision: int = FP64, device: int = 0, csv_path: str = None,
             echo: bool = False, export_dir: str = None, export_interval: int = 0):
    """The reference driver loop on the device for a params.json path, JSON text, dict or Params.

    Returns the logged energy rows as an array [k, 8]: step, KE, PE_trap, PE_coul, total, n_active,
    fractional drift (of e |U0|), absolute drift.  csv_path receives the reference's energy CSV;
    with export_dir and export_interval > 0 the ParaView files of every export_interval-th step (and
    step 0) go to export_dir (an existing directory)."""
    if isinstance(config, Params):
        p = config
    elif isinstance(config, str) and not config.lstrip().startswith("{"):
        p = load_params(config)
    else:
        p = parse_params(config)
    opt = RunOptions(device, int(precision), int(rng_seed), int(echo), 0, int(export_interval),
                     None if export_dir is None else export_dir.encode())
    cap = int(min(p.num_steps, 1 << 22)) + 2
    rows = (EnergyRow * cap)()
    n = ctypes.c_int64()
    _check(_bind().am_simulate(ctypes.byref(p), ctypes.byref(opt), None if csv_path is None else csv_path.encode(),
                               ctypes.cast(rows, _vp), cap, ctypes.byref(n)))
    out = np.zeros((n.value, 8))
    for k in range(n.value):
        r = rows[k]
        out[k] = (r.step, r.kinetic, r.trap_potential, r.coulomb_potential, r.total, r.num_active_particles,
                  r.fractional_drift, r.absolute_drift)
    return out


def export_vtu(store, path: str):
    """The store's agents still in the simulation as a ParaView .vtu file (am_export_vtu)."""
    _check(_bind().am_export_vtu(store.h, path.encode()))
