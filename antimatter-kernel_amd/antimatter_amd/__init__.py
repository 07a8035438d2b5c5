"""Python binding of libantimatter_amd.so (the MI355X hot path of antimatter-kernel).

This module is a thin ctypes layer over the C ABI declared in
``include/antimatter_amd/am_api.h``; every computation runs in the HIP library.  It
raises ``AmError`` when the library returns an error and ``RuntimeError`` when the
library is missing: there is no CPU fallback.

Names follow the reference's free functions and behaviours
(src/interaction/*.h, src/forces/*.h, src/analysis/energy_analysis.h).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import weakref

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.dirname(_PKG_DIR)
LIB_PATH = os.path.join(LIB_DIR, "libantimatter_amd.so")
REPO_ROOT = os.path.dirname(LIB_DIR)
HEADER_PATH = os.path.join(REPO_ROOT, "include", "antimatter_amd", "am_api.h")

AM_OK = 0
FP64 = 0
FP32 = 1

SPECIES = {"antiproton": 0, "positron": 1, "antihydrogen": 2, "electron": 3, "proton": 4, "antihelium": 5}
METHODS = {"standard": 0, "barnes-hut": 1, "fast-multipole": 2, "all-pairs": 3}

_d = ctypes.c_double
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_int = ctypes.c_int
_vp = ctypes.c_void_p


class AmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"antimatter_amd error {code}: {msg}")
        self.code = code


class TrapField(ctypes.Structure):
    """PenningTrapField (src/fields/penning_trap.h:27-69)."""
    _fields_ = [("b0", _d), ("trap_voltage", _d), ("d", _d)]


class EnergyBreakdown(ctypes.Structure):
    _fields_ = [("kinetic", _d), ("trap_potential", _d), ("energy_potential", _d), ("num_active_particles", _i64)]

    def total(self) -> float:
        return self.kinetic + self.trap_potential + self.energy_potential


class EnergyHistory(ctypes.Structure):
    _fields_ = [("has_previous", _i32), ("prev_trap_potential", _d), ("prev_coulomb_potential", _d)]


class EnergySettings(ctypes.Structure):
    _fields_ = [("method", _i32), ("cutoff_radius", _d), ("barnes_hut_theta", _d), ("fast_multipole_theta", _d),
                ("fast_multipole_order", _d)]


# Exported C-ABI symbols and their signatures (restype, argtypes).
SIGNATURES = {
    "am_last_error": (ctypes.c_char_p, []),
    "am_version": (_int, []),
    "am_trap_make_from_geometry": (_int, [_d, _d, _d, _d, _vp]),
    "am_cooling_time": (_d, [_d, _d, _d]),
    "am_species_properties": (_int, [_int, _vp, _vp]),
    "am_is_antiparticle_pair": (_int, [_int, _int]),
    "am_set_particle_positions": (_int, [_vp, _vp, _int, _vp]),
    "am_set_particle_velocities": (_int, [_vp, _vp, _int, _vp]),
    "am_derive_dt": (_d, [_vp, _d]),
    "am_ctx_create": (_int, [_int, _vp]),
    "am_ctx_destroy": (_int, [_vp]),
    "am_ctx_synchronize": (_int, [_vp]),
    "am_ctx_enable_timing": (_int, [_vp, _int]),
    "am_ctx_kernel_time": (_int, [_vp, ctypes.c_char_p, _vp, _vp]),
    "am_ctx_reset_timing": (_int, [_vp]),
    "am_store_create": (_int, [_vp, _i64, _vp]),
    "am_store_destroy": (_int, [_vp]),
    "am_store_upload": (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "am_store_download": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "am_store_set_coulomb_force": (_int, [_vp, _vp]),
    "am_store_count": (_int, [_vp, _vp, _vp]),
    "am_store_generate_cloud": (_int, [_vp, _vp, _vp, _vp, _u64]),
    "am_store_download_species": (_int, [_vp, _vp]),
    "am_compute_standard_coulomb_forces": (_int, [_vp, _d, _int]),
    "am_compute_allpairs_coulomb_forces": (_int, [_vp, _int]),
    "am_allpairs_forces_noalloc": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "am_compute_barnes_hut_coulomb_forces": (_int, [_vp, _d, _int]),
    "am_barnes_hut_forces_counted": (_int, [_vp, _d, _int, _vp]),
    "am_barnes_hut_evaluate": (_int, [_vp, _d, _int, _i64, _vp, _vp, _vp, _vp, _vp]),
    "am_bh_tree_stats": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "am_bh_tree_download": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "am_compute_fast_multipole_coulomb_forces": (_int, [_vp, _int, _d]),
    "am_fmm_stats": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "am_fmm_m2l_stats": (_int, [_vp, _vp, _vp]),
    "am_compute_fast_multipole_coulomb_forces_ex": (_int, [_vp, _int, _d, _int]),
    "am_lorentz_force": (_int, [_vp, _vp, _d]),
    "am_cyclotron_cooling": (_int, [_vp, _vp, _d, _d, _d, _u64, _u64]),
    "am_push": (_int, [_vp, _vp, _d, _d, _d, _u64, _u64]),
    "am_annihilation": (_int, [_vp, _d, _d, _u64, _u64, _vp]),
    "am_compute_total_energy": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "am_compute_total_energy_ex": (_int, [_vp, _vp, _vp, _vp, _int, _vp]),
    "am_ensemble_create": (_int, [_vp, _i32, _vp, _i64, _vp]),
    "am_ensemble_destroy": (_int, [_vp]),
    "am_ensemble_set_chunk": (_int, [_vp, _i32]),
    "am_ensemble_autotune": (_int, [_vp, _vp, _i32, _vp, _vp]),
    "am_ensemble_run": (_int, [_vp, _i32, _u64, _i32]),
    "am_ensemble_stats": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "am_mlp_create": (_int, [_vp, _u64, _vp]),
    "am_mlp_destroy": (_int, [_vp]),
    "am_mlp_num_params": (_int, []),
    "am_mlp_set_params": (_int, [_vp, _vp]),
    "am_mlp_get_params": (_int, [_vp, _vp]),
    "am_mlp_set_normalization": (_int, [_vp, _vp, _vp, _vp]),
    "am_mlp_set_baseline": (_int, [_vp, _int]),
    "am_mlp_quadrupole_baseline": (_int, [_vp, _i64, _vp]),
    "am_mlp_gradients": (_int, [_vp, _vp, _vp, _i32, _vp, _vp]),
    "am_mlp_train": (_int, [_vp, _vp, _vp, _i64, _i32, _i32, ctypes.c_float, _u64, _vp]),
    "am_mlp_infer": (_int, [_vp, _vp, _i32, _vp]),
    "am_bh_learned_training_data": (_int, [_vp, _d, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "am_compute_barnes_hut_learned_forces": (_int, [_vp, _d, _vp, _i32, ctypes.c_float, ctypes.c_float, _vp]),
    "am_debug_exclusive_scan": (_int, [_vp, _vp, _i64, _vp]),
    "am_debug_radix_sort": (_int, [_vp, _vp, _vp, _i64, _int]),
    "am_debug_radix_sort_keys": (_int, [_vp, _vp, _i64, _int, _int]),
    "am_debug_fmm_m2l_options": (_int, [_vp, _int, _i64]),
}

_lib = None


def build(force: bool = True) -> None:
    """Compile libantimatter_amd.so for gfx950 with the package Makefile (hipcc)."""
    args = ["make", "-s", "-j8", "-C", LIB_DIR]
    if force:
        subprocess.run(["make", "-s", "-C", LIB_DIR, "clean"], check=True)
    subprocess.run(args, check=True)


def lib():
    """Load the HIP library (raises if it has not been built: no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the HIP path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != AM_OK:
        raise AmError(rc, lib().am_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data


def make_trap(b0, trap_voltage, z0, rho0) -> TrapField:
    """PenningTrapField::MakeFromGeometry (src/fields/penning_trap.cc:15-20)."""
    t = TrapField()
    _check(lib().am_trap_make_from_geometry(b0, trap_voltage, z0, rho0, ctypes.byref(t)))
    return t


def cooling_time(charge, mass, b_mag) -> float:
    return float(lib().am_cooling_time(charge, mass, b_mag))


def species_properties(species: int):
    m, q = ctypes.c_double(), ctypes.c_double()
    _check(lib().am_species_properties(int(species), ctypes.byref(m), ctypes.byref(q)))
    return m.value, q.value


def is_antiparticle_pair(a: int, b: int) -> bool:
    return bool(lib().am_is_antiparticle_pair(int(a), int(b)))


class Context:
    """One GPU device and one HIP stream (am_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().am_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # stores made from this context (closed first)

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            lib().am_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().am_ctx_synchronize(self.h))

    def exclusive_scan(self, counts):
        a = np.ascontiguousarray(counts, dtype=np.uint32)
        out = np.zeros(len(a) + 1, dtype=np.uint32)
        _check(lib().am_debug_exclusive_scan(self.h, _p(a), len(a), _p(out)))
        return out

    def radix_sort(self, keys, vals, key_bits=64):
        k = np.array(keys, dtype=np.uint64)
        v = np.array(vals, dtype=np.uint32)
        _check(lib().am_debug_radix_sort(self.h, _p(k), _p(v), len(k), int(key_bits)))
        return k, v

    def radix_sort_keys(self, keys, bit0, key_bits):
        """Stable sort of 64-bit keys on bits [bit0, bit0 + key_bits) alone (device)."""
        k = np.array(keys, dtype=np.uint64)
        _check(lib().am_debug_radix_sort_keys(self.h, _p(k), len(k), int(bit0), int(key_bits)))
        return k

    def fmm_m2l_options(self, classes: bool = True, max_unclassed: int = 1 << 20):
        """FMM M2L diagnostics: classes=False gives every M2L pair its own tensor; past max_unclassed
        unclassed pairs (always, when negative) FP32 FMM calls run the FP64 contraction."""
        _check(lib().am_debug_fmm_m2l_options(self.h, int(bool(classes)), int(max_unclassed)))

    def enable_timing(self, on: bool = True):
        _check(lib().am_ctx_enable_timing(self.h, int(on)))

    def reset_timing(self):
        _check(lib().am_ctx_reset_timing(self.h))

    def kernel_time(self, name: str):
        """(total milliseconds, launches) of a timed kernel family."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        _check(lib().am_ctx_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


class Store:
    """SoA particle store in HBM (am_store): the agents of the reference's ResourceManager."""

    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().am_store_create(ctx.h, int(capacity), ctypes.byref(h)))
        self.h = h
        self.capacity = int(capacity)
        self._children = weakref.WeakSet()  # ensembles over this store (closed first)
        ctx._children.add(self)

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            if self.ctx.h:
                lib().am_store_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos, vel, species, charge=None, mass=None, annihilated=None, trap=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        species = np.ascontiguousarray(species, dtype=np.uint8)
        n = len(species)
        if pos.size != 3 * n or vel.size != 3 * n:
            raise ValueError("pos/vel must hold 3 values per particle")
        c = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        a = None if annihilated is None else np.ascontiguousarray(annihilated, dtype=np.uint8)
        t = None if trap is None else np.ascontiguousarray(trap, dtype=np.uint32)
        _check(lib().am_store_upload(self.h, n, _p(pos), _p(vel), _p(species), _p(c), _p(m), _p(a), _p(t)))
        self.n = n

    def generate_cloud(self, populations: dict, temperatures: dict, radii, seed: int):
        """Seeded cloud drawn on the device (am_store_generate_cloud): the reference's distributions,
        not its draws; for clouds of millions of particles."""
        c = np.zeros(6, np.uint64)
        t = np.zeros(6)
        for name, cnt in populations.items():
            c[SPECIES[name] if isinstance(name, str) else int(name)] = int(cnt)
        for name, T in temperatures.items():
            t[SPECIES[name] if isinstance(name, str) else int(name)] = float(T)
        r = np.ascontiguousarray(radii, dtype=np.float64)
        _check(lib().am_store_generate_cloud(self.h, _p(c), _p(r), _p(t), int(seed)))
        self.n = int(c.sum())

    def species(self):
        n = self.count()[0]
        sp = np.zeros(n, np.uint8)
        _check(lib().am_store_download_species(self.h, _p(sp)))
        return sp

    def download(self):
        n = ctypes.c_int64()
        _check(lib().am_store_count(self.h, ctypes.byref(n), None))
        n = n.value
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        force = np.zeros((n, 3))
        ann = np.zeros(n, dtype=np.uint8)
        stag = np.zeros(n, dtype=np.uint8)
        _check(lib().am_store_download(self.h, _p(pos), _p(vel), _p(force), _p(ann), _p(stag)))
        return dict(pos=pos, vel=vel, force=force, annihilated=ann, staggered=stag)

    def forces(self):
        n = self.count()[0]
        f = np.zeros((n, 3))
        _check(lib().am_store_download(self.h, None, None, _p(f), None, None))
        return f

    def set_coulomb_force(self, force):
        f = np.ascontiguousarray(force, dtype=np.float64)
        _check(lib().am_store_set_coulomb_force(self.h, _p(f)))

    def count(self):
        n = ctypes.c_int64()
        a = ctypes.c_int64()
        _check(lib().am_store_count(self.h, ctypes.byref(n), ctypes.byref(a)))
        return n.value, a.value

    # ---- solvers (ComputeStandard/BarnesHut/FastMultipoleCoulombForces) ----
    def compute_standard_coulomb_forces(self, cutoff_radius: float, precision: int = FP64):
        _check(lib().am_compute_standard_coulomb_forces(self.h, cutoff_radius, precision))

    def compute_allpairs_coulomb_forces(self, precision: int = FP32):
        _check(lib().am_compute_allpairs_coulomb_forces(self.h, precision))

    def compute_barnes_hut_coulomb_forces(self, theta: float, precision: int = FP32):
        _check(lib().am_compute_barnes_hut_coulomb_forces(self.h, theta, precision))

    def barnes_hut_forces_counted(self, theta: float, precision: int = FP32):
        n = self.count()[0]
        c = np.zeros(n, dtype=np.int64)
        _check(lib().am_barnes_hut_forces_counted(self.h, theta, precision, _p(c)))
        return c

    def barnes_hut_evaluate(self, theta: float, points, charges=None, self_index=None, precision: int = FP64,
                            want_force: bool = True, want_potential: bool = True):
        """BarnesHutTree::ComputeForce / ComputePotential at given points over the store's particles:
        (forces [k, 3] in N or None, potentials [k] in C/m or None)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        k = len(pts)
        q = None if charges is None else np.ascontiguousarray(charges, dtype=np.float64)
        si = None if self_index is None else np.ascontiguousarray(self_index, dtype=np.int64)
        f = np.zeros((k, 3)) if want_force else None
        phi = np.zeros(k) if want_potential else None
        _check(lib().am_barnes_hut_evaluate(self.h, theta, int(precision), k, _p(pts), _p(q), _p(si), _p(f), _p(phi)))
        return f, phi

    def bh_tree(self):
        """Tree of the most recent Barnes-Hut build (reference DFS preorder)."""
        m, nn, md, ov = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        _check(lib().am_bh_tree_stats(self.h, ctypes.byref(m), ctypes.byref(nn), ctypes.byref(md),
                                      ctypes.byref(ov)))
        order = np.zeros(m.value, np.uint32)
        b = np.zeros(nn.value, np.int32)
        e = np.zeros(nn.value, np.int32)
        leaf = np.zeros(nn.value, np.uint8)
        qp = np.zeros(nn.value)
        qn = np.zeros(nn.value)
        _check(lib().am_bh_tree_download(self.h, _p(order), _p(b), _p(e), _p(leaf), _p(qp), _p(qn)))
        return dict(order=order, begin=b, end=e, leaf=leaf, qp=qp, qn=qn, max_depth=md.value,
                    deep_overflow=ov.value)

    def compute_fast_multipole_coulomb_forces(self, order: int, theta: float, precision: int = FP64):
        if precision == FP64:
            _check(lib().am_compute_fast_multipole_coulomb_forces(self.h, int(order), theta))
        else:
            _check(lib().am_compute_fast_multipole_coulomb_forces_ex(self.h, int(order), theta, precision))

    def fmm_stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        d = ctypes.c_int32()
        c = ctypes.c_int64()
        _check(lib().am_fmm_stats(self.h, *[ctypes.byref(x) for x in v], ctypes.byref(d), ctypes.byref(c)))
        u, it = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().am_fmm_m2l_stats(self.h, ctypes.byref(u), ctypes.byref(it)))
        return dict(sources=v[0].value, cells=v[1].value, m2l_pairs=v[2].value, p2p_cell_pairs=v[3].value,
                    max_depth=d.value, m2l_classes=c.value, m2l_unclassed=u.value, m2l_items=it.value)

    # ---- behaviours ----
    def lorentz_force(self, field: TrapField, dt: float):
        _check(lib().am_lorentz_force(self.h, ctypes.byref(field), dt))

    def cyclotron_cooling(self, field: TrapField, tau_scale, bath_temp, dt, seed, step):
        _check(lib().am_cyclotron_cooling(self.h, ctypes.byref(field), tau_scale, bath_temp, dt, seed, step))

    def push(self, field: TrapField, dt, tau_scale=0.0, bath_temp=0.0, seed=0, step=0):
        _check(lib().am_push(self.h, ctypes.byref(field), dt, tau_scale, bath_temp, seed, step))

    def annihilation(self, contact_radius, probability, seed, step) -> int:
        k = ctypes.c_int64()
        _check(lib().am_annihilation(self.h, contact_radius, probability, seed, step, ctypes.byref(k)))
        return k.value

    def compute_total_energy(self, field: TrapField, method="standard", cutoff_radius=0.0, barnes_hut_theta=0.5,
                             fast_multipole_theta=0.5, fast_multipole_order=4, history: EnergyHistory = None,
                             precision: int = FP64):
        st = EnergySettings(METHODS[method] if isinstance(method, str) else int(method), cutoff_radius,
                            barnes_hut_theta, fast_multipole_theta, float(fast_multipole_order))
        out = EnergyBreakdown()
        _check(lib().am_compute_total_energy_ex(self.h, ctypes.byref(field), ctypes.byref(st),
                                                None if history is None else ctypes.byref(history), int(precision),
                                                ctypes.byref(out)))
        return out


from .driver import simulate, derive_dt, initial_conditions  # noqa: E402,F401
from .ensemble import Ensemble, TrapDesc, make_config5, trap_seed  # noqa: E402,F401
