"""ORACLE — test infrastructure only.

ctypes binding of ``oracle/liboracle.so``, the FP64 CPU restatement of the
antimatter-kernel hot path (see ``cpu_ref.cpp`` for the per-function reference
citations and what pins each function).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import this package, as the checker or
as the timed CPU baseline.  The product library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SPECIES = {"antiproton": 0, "positron": 1, "antihydrogen": 2, "electron": 3, "proton": 4, "antihelium": 5}

# src/utils/constants.h:11-38
E_CHARGE = 1.602176634e-19
M_E = 9.1093837015e-31
M_P = 1.67262192369e-27
K_COULOMB = 8.9875517923e9
K_BOLTZMANN = 1.380649e-23

_d = ctypes.c_double
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_vp = ctypes.c_void_p


def build() -> None:
    """Compile liboracle.so with the committed Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_philox4x32_10": (None, [_vp, _vp, _vp]),
            "orc_annihilation_uniform": (_d, [_u64, _u64, _u32, _u32]),
            "orc_cooling_normals": (None, [_u64, _u64, _u32, _vp]),
            "orc_set_num_threads": (None, [ctypes.c_int]),
            "orc_max_threads": (ctypes.c_int, []),
            "orc_species_props": (None, [ctypes.c_int, _vp, _vp]),
            "orc_derive_dt": (_d, [_vp, _d]),
            "orc_set_particle_positions": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
            "orc_set_particle_velocities": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
            "orc_trap_characteristic_dimension": (None, [_d, _d, _vp]),
            "orc_electric_field": (None, [_d, _d, _i64, _vp, _vp]),
            "orc_potential": (None, [_d, _d, _i64, _vp, _vp]),
            "orc_cooling_time": (_d, [_d, _d, _d]),
            "orc_lorentz_push": (None, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, _d]),
            "orc_cyclotron_cooling": (None, [_i64, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _u64, _u64]),
            "orc_standard_forces": (None, [_i64, _vp, _vp, _vp, _d, _vp]),
            "orc_allpairs_forces": (None, [_i64, _vp, _vp, _vp, _vp]),
            "orc_allpairs_forces_subset": (None, [_i64, _vp, _vp, _vp, _i64, _vp, _vp]),
            "orc_allpairs_potential_energy": (_d, [_i64, _vp, _vp, _vp]),
            "orc_standard_energy": (_d, [_i64, _vp, _vp, _vp, _d]),
            "orc_bh_forces": (None, [_i64, _vp, _vp, _vp, _d, _i64, _vp, _vp]),
            "orc_bh_potential_energy": (_d, [_i64, _vp, _vp, _vp, _d]),
            "orc_bh_potentials": (None, [_i64, _vp, _vp, _vp, _d, _i64, _vp]),
            "orc_bh_evaluate": (None, [_i64, _vp, _vp, _vp, _d, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
            "orc_bh_tree": (_i64, [_i64, _vp, _vp, _i64, _i64] + [_vp] * 12),
            "orc_fmm_forces": (None, [_i64, _vp, _vp, _vp, ctypes.c_int, _d, _i64, _vp, _vp]),
            "orc_fmm_potential_energy": (_d, [_i64, _vp, _vp, _vp, ctypes.c_int, _d]),
            "orc_fmm_derivatives": (ctypes.c_int, [ctypes.c_int, _d, _d, _d, _vp, _vp]),
            "orc_annihilate": (_i64, [_i64, _vp, _vp, _vp, _d, _d, _u64, _u64, _vp]),
            "orc_energy": (_i64, [_i64, _vp, _vp, _vp, _vp, _vp, _d, _d, ctypes.c_int, _d, _d, _d, _d, _vp]),
            "orc_simulate": (_i64, [_vp, _vp, _i64, _vp, _vp, _vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def annihilation_uniform(seed, step, i, j) -> float:
    return float(lib().orc_annihilation_uniform(seed, step, i, j))


def cooling_normals(seed, step, i):
    out = np.zeros(3)
    lib().orc_cooling_normals(seed, step, i, _p(out))
    return out


def species_props(species: int):
    m = ctypes.c_double()
    q = ctypes.c_double()
    lib().orc_species_props(int(species), ctypes.byref(m), ctypes.byref(q))
    return m.value, q.value


def counts_array(populations: dict):
    c = np.zeros(6, dtype=np.uint64)
    listed = np.zeros(6, dtype=np.int32)
    for name, cnt in populations.items():
        s = SPECIES[name] if isinstance(name, str) else int(name)
        c[s] = int(cnt)
        listed[s] = 1
    return c, listed


def derive_dt(populations: dict, b0: float) -> float:
    _, listed = counts_array(populations)
    return float(lib().orc_derive_dt(_p(listed), b0))


def set_particle_positions(populations: dict, radii, seed: int):
    c, _ = counts_array(populations)
    n = int(c.sum())
    r = np.asarray(radii, dtype=np.float64)
    out = np.zeros((n, 3))
    lib().orc_set_particle_positions(_p(c), _p(r), int(seed), _p(out))
    return out


def set_particle_velocities(populations: dict, temperatures: dict, seed: int):
    c, _ = counts_array(populations)
    t = np.zeros(6)
    for name, T in temperatures.items():
        t[SPECIES[name] if isinstance(name, str) else int(name)] = float(T)
    n = int(c.sum())
    out = np.zeros((n, 3))
    lib().orc_set_particle_velocities(_p(c), _p(t), int(seed), _p(out))
    return out


def species_array(populations: dict):
    c, _ = counts_array(populations)
    return np.concatenate([np.full(int(c[s]), s, dtype=np.uint8) for s in range(6)])


def mass_charge(species):
    m = np.zeros(len(species))
    q = np.zeros(len(species))
    for s in range(6):
        ms, qs = species_props(s)
        m[species == s] = ms
        q[species == s] = qs
    return m, q


def characteristic_dimension(z0, rho0) -> float:
    d = ctypes.c_double()
    lib().orc_trap_characteristic_dimension(z0, rho0, ctypes.byref(d))
    return d.value


def electric_field(u0, d, pos):
    pos = _c(pos, np.float64)
    e = np.zeros_like(pos)
    lib().orc_electric_field(u0, d, len(pos), _p(pos), _p(e))
    return e


def potential(u0, d, pos):
    pos = _c(pos, np.float64)
    phi = np.zeros(len(pos))
    lib().orc_potential(u0, d, len(pos), _p(pos), _p(phi))
    return phi


def cooling_time(q, m, b) -> float:
    return float(lib().orc_cooling_time(q, m, b))


def lorentz_push(pos, vel, fcoul, q, m, annihilated, staggered, b0, u0, d, dt):
    """In-place push of pos/vel; fcoul is zeroed; staggered flags updated."""
    lib().orc_lorentz_push(len(q), _p(pos), _p(vel), _p(fcoul), _p(q), _p(m), _p(annihilated), _p(staggered),
                           b0, u0, d, dt)


def cyclotron_cooling(vel, q, m, annihilated, b0, tau_scale, bath_temp, dt, seed, step):
    lib().orc_cyclotron_cooling(len(q), _p(vel), _p(q), _p(m), _p(annihilated), b0, tau_scale, bath_temp, dt,
                                seed, step)


def standard_forces(pos, q, annihilated, rc):
    pos = _c(pos, np.float64)
    f = np.zeros_like(pos)
    lib().orc_standard_forces(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), rc, _p(f))
    return f


def allpairs_forces(pos, q, annihilated=None):
    pos = _c(pos, np.float64)
    f = np.zeros_like(pos)
    lib().orc_allpairs_forces(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), _p(f))
    return f


def allpairs_forces_subset(pos, q, annihilated, targets):
    pos = _c(pos, np.float64)
    t = _c(targets, np.int64)
    f = np.zeros((len(t), 3))
    lib().orc_allpairs_forces_subset(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), len(t),
                                     _p(t), _p(f))
    return f


def allpairs_potential_energy(pos, q, annihilated=None) -> float:
    return float(lib().orc_allpairs_potential_energy(len(q), _p(_c(pos, np.float64)), _p(_c(q, np.float64)),
                                                     _p(_c(annihilated, np.uint8))))


def standard_energy(pos, q, annihilated, rc) -> float:
    return float(lib().orc_standard_energy(len(q), _p(_c(pos, np.float64)), _p(_c(q, np.float64)),
                                           _p(_c(annihilated, np.uint8)), rc))


def bh_forces(pos, q, annihilated, theta, leaf_cap=8, want_counts=False):
    pos = _c(pos, np.float64)
    f = np.zeros_like(pos)
    cnt = np.zeros(len(q), dtype=np.int64) if want_counts else None
    lib().orc_bh_forces(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), theta, leaf_cap,
                        _p(f), _p(cnt))
    return (f, cnt) if want_counts else f


def bh_potential_energy(pos, q, annihilated, theta) -> float:
    return float(lib().orc_bh_potential_energy(len(q), _p(_c(pos, np.float64)), _p(_c(q, np.float64)),
                                               _p(_c(annihilated, np.uint8)), theta))


def bh_potentials(pos, q, annihilated, theta, leaf_cap=8):
    """Per-particle BarnesHutTree::ComputePotential(x_i, i) in C/m (no Coulomb constant)."""
    pos = _c(pos, np.float64)
    phi = np.zeros(len(q))
    lib().orc_bh_potentials(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), theta,
                            leaf_cap, _p(phi))
    return phi


def bh_evaluate(pos, q, annihilated, theta, points, charges=None, self_index=None, leaf_cap=8):
    """BarnesHutTree::ComputeForce / ComputePotential at given points: (forces [k, 3], phi [k] in C/m)."""
    pos = _c(pos, np.float64)
    pts = _c(points, np.float64).reshape(-1, 3)
    k = len(pts)
    f = np.zeros((k, 3))
    phi = np.zeros(k)
    lib().orc_bh_evaluate(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), theta, leaf_cap, k,
                          _p(pts), _p(_c(charges, np.float64)), _p(_c(self_index, np.int64)), _p(f), _p(phi))
    return f, phi


def bh_tree(pos, q, leaf_cap=8):
    """Reference-order octree (DFS preorder) as a dict of arrays."""
    pos = _c(pos, np.float64)
    q = _c(q, np.float64)
    n = len(q)
    order = np.zeros(n, dtype=np.int64)
    nn = lib().orc_bh_tree(n, _p(pos), _p(q), leaf_cap, 0, _p(order), *([None] * 11))
    t = dict(order=order, begin=np.zeros(nn, np.int64), end=np.zeros(nn, np.int64), depth=np.zeros(nn, np.int32),
             leaf=np.zeros(nn, np.uint8), center=np.zeros((nn, 3)), half=np.zeros(nn), qp=np.zeros(nn),
             qn=np.zeros(nn), cp=np.zeros((nn, 3)), cn=np.zeros((nn, 3)), child=np.zeros((nn, 8), np.int32))
    lib().orc_bh_tree(n, _p(pos), _p(q), leaf_cap, nn, _p(order), _p(t["begin"]), _p(t["end"]), _p(t["depth"]),
                      _p(t["leaf"]), _p(t["center"]), _p(t["half"]), _p(t["qp"]), _p(t["qn"]), _p(t["cp"]),
                      _p(t["cn"]), _p(t["child"]))
    return t


def fmm_forces(pos, q, annihilated, order, theta, leaf_cap=32, want_stats=False):
    pos = _c(pos, np.float64)
    f = np.zeros_like(pos)
    st = np.zeros(4, dtype=np.int64)
    lib().orc_fmm_forces(len(q), _p(pos), _p(_c(q, np.float64)), _p(_c(annihilated, np.uint8)), int(order), theta,
                         leaf_cap, _p(f), _p(st))
    return (f, dict(cells=int(st[0]), leaves=int(st[1]), m2l_pairs=int(st[2]), p2p_pairs=int(st[3]))) \
        if want_stats else f


def fmm_potential_energy(pos, q, annihilated, order, theta) -> float:
    return float(lib().orc_fmm_potential_energy(len(q), _p(_c(pos, np.float64)), _p(_c(q, np.float64)),
                                                _p(_c(annihilated, np.uint8)), int(order), theta))


def fmm_derivatives(order, r):
    nd = (2 * order + 1) * (2 * order + 2) * (2 * order + 3) // 6
    D = np.zeros(nd)
    multi = np.zeros((nd, 3), dtype=np.int32)
    lib().orc_fmm_derivatives(int(order), float(r[0]), float(r[1]), float(r[2]), _p(D), _p(multi))
    return D, multi


def annihilate(pos, species, annihilated, ra, p, seed, step):
    """In-place canonical annihilation; returns (count, pairs array)."""
    n = len(species)
    pairs = np.zeros((n, 2), dtype=np.int64)
    cnt = lib().orc_annihilate(n, _p(pos), _p(species), _p(annihilated), ra, p, seed, step, _p(pairs))
    return int(cnt), pairs[: cnt // 2].copy()


def energy(pos, vel, q, m, annihilated, u0, d, method=0, rc=0.0, bh_theta=0.5, fmm_theta=0.5, fmm_order=4):
    out = np.zeros(3)
    act = lib().orc_energy(len(q), _p(_c(pos, np.float64)), _p(_c(vel, np.float64)), _p(_c(q, np.float64)),
                           _p(_c(m, np.float64)), _p(_c(annihilated, np.uint8)), u0, d, int(method), rc, bh_theta,
                           fmm_theta, float(fmm_order), _p(out))
    return dict(kinetic=out[0], trap_potential=out[1], energy_potential=out[2], num_active_particles=int(act))


class _SimConfig(ctypes.Structure):
    _fields_ = [("counts", _u64 * 6), ("listed", _i32 * 6), ("temps", _d * 6), ("radii", _d * 3), ("seed", _i32),
                ("b0", _d), ("u0", _d), ("z0", _d), ("rho0", _d), ("num_steps", _i64), ("cooling_scale", _d),
                ("bath_temp", _d), ("coulomb_enabled", _i32), ("coulomb_method", _i32), ("energy_method", _i32),
                ("cutoff_radius", _d), ("bh_theta", _d), ("fmm_theta", _d), ("fmm_order", _d),
                ("annihilation_enabled", _i32), ("contact_radius", _d), ("annihilation_probability", _d),
                ("rng_seed", _u64)]


_METHODS = {"standard": 0, "barnes-hut": 1, "fast-multipole": 2, "all-pairs": 3}


def simulate(params: dict, rng_seed: int = 42):
    """Run the reference driver loop (antimatter-kernel.cc:86-197) on the oracle.

    ``params`` uses the params.json keys.  Returns (rows[k, 8], final dict).
    """
    c = _SimConfig()
    counts, listed = counts_array(params["particle_populations"])
    n = int(counts.sum())
    for s in range(6):
        c.counts[s] = int(counts[s])
        c.listed[s] = int(listed[s])
    for name, T in params.get("particle_temperatures", {}).items():
        c.temps[SPECIES[name]] = float(T)
    r = params.get("initial_cloud_radii", {})
    c.radii[0] = r.get("radius_x", 1e-4)
    c.radii[1] = r.get("radius_y", 1e-4)
    c.radii[2] = r.get("radius_z", 1e-4)
    c.seed = int(params.get("seed", 42))
    c.b0 = float(params.get("b0", 1.0))
    c.u0 = float(params.get("trap_voltage", 10.0))
    c.z0 = float(params.get("z0", 0.01))
    c.rho0 = float(params.get("rho0", 0.01))
    c.num_steps = int(params.get("num_steps", 5000))
    c.cooling_scale = float(params.get("cooling_scale", 10.0))
    c.bath_temp = float(params.get("bath_temperature", 0.5))
    c.coulomb_enabled = int(bool(params.get("coulomb_forces_enabled", True)))
    c.coulomb_method = _METHODS[params.get("coulomb_computation_method", "standard")]
    c.energy_method = _METHODS[params.get("energy_computation_method", "standard")]
    c.cutoff_radius = float(params.get("cutoff_radius", 0.0))
    c.bh_theta = float(params.get("barnes_hut_theta", 0.5))
    c.fmm_theta = float(params.get("fast_multipole_theta", 0.5))
    c.fmm_order = float(params.get("fast_multipole_order", 4))
    c.annihilation_enabled = int(bool(params.get("annihilation_enabled", False)))
    c.contact_radius = float(params.get("contact_radius", 1e-9))
    c.annihilation_probability = float(params.get("annihilation_probability", 0.1))
    c.rng_seed = int(rng_seed)
    max_rows = c.num_steps + 1
    rows = np.zeros((max_rows, 8))
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    ann = np.zeros(n, dtype=np.uint8)
    k = lib().orc_simulate(ctypes.byref(c), _p(rows), max_rows, _p(pos), _p(vel), _p(ann))
    if k < 0:
        raise ValueError("invalid simulation config")
    return rows[:k].copy(), dict(pos=pos, vel=vel, annihilated=ann)
