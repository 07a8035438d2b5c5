"""Initial conditions, the dt rule and the driver loop of ``Simulate`` (src/antimatter-kernel.cc:42-217).

The driver loop itself is the library's C++ ``am_simulate`` (antimatter-kernel_amd/host/simulate.cpp):
per step the Coulomb solver, the fused LorentzForce + CyclotronCooling push, AnnihilationInteraction
in canonical order on the pushed positions, and the time-centred energy at log steps and the step
before them (antimatter-kernel.cc:153-197).  ``simulate`` here takes the params.json keys as a dict
(or a path or JSON text) and returns the logged rows.
"""
from __future__ import annotations

import numpy as np

from . import FP64, SPECIES, _check, _p, lib

E_CHARGE = 1.602176634e-19


def counts_array(populations: dict):
    c = np.zeros(6, dtype=np.uint64)
    listed = np.zeros(6, dtype=np.int32)
    for name, cnt in populations.items():
        s = SPECIES[name] if isinstance(name, str) else int(name)
        c[s] = int(cnt)
        listed[s] = 1
    return c, listed


def derive_dt(populations: dict, b0: float) -> float:
    """PhysicsParams dt rule (src/params/physics_params.cc:60-74)."""
    _, listed = counts_array(populations)
    return float(lib().am_derive_dt(_p(listed), b0))


def initial_conditions(populations: dict, temperatures: dict, radii, seed: int):
    """Positions, velocities and species of the reference IC generators (penning_trap.cc:51-161)."""
    c, _ = counts_array(populations)
    n = int(c.sum())
    t = np.zeros(6)
    for name, T in temperatures.items():
        t[SPECIES[name] if isinstance(name, str) else int(name)] = float(T)
    r = np.asarray(radii, dtype=np.float64)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    _check(lib().am_set_particle_positions(_p(c), _p(r), int(seed), _p(pos)))
    _check(lib().am_set_particle_velocities(_p(c), _p(t), int(seed), _p(vel)))
    species = np.concatenate([np.full(int(c[s]), s, dtype=np.uint8) for s in range(6)])
    return pos, vel, species


def simulate(params, rng_seed: int = 42, precision: int = FP64, device: int = 0, csv_path: str = None,
             export_dir: str = None, export_interval: int = 0):
    """Run the reference driver loop on the GPU (the library's am_simulate).

    ``params``: a dict of params.json keys, a path to a params.json, or JSON text; it is read with the
    reference's conversions and must pass the reference's Validate checks.  Returns rows[k, 8]:
    step, KE, PE_trap, PE_coul, total, n_active, fractional drift (of e*|U0|), absolute drift.
    With ``export_dir`` and ``export_interval`` > 0, the ParaView files of every export_interval-th step
    (the reference's BioDynaMo export, bdm.toml:6-13) go to that directory.
    """
    from .params import simulate as native
    return native(params, rng_seed=rng_seed, precision=precision, device=device, csv_path=csv_path,
                  export_dir=export_dir, export_interval=export_interval)
