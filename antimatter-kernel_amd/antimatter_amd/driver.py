"""Host driver of the simulation loop, mirroring ``Simulate`` (src/antimatter-kernel.cc:42-217).

Step order per step (the reference's contract, antimatter-kernel.cc:153-197):
Coulomb solver -> LorentzForce + CyclotronCooling (fused push) -> AnnihilationInteraction
(canonical order, on post-push positions) -> energy at log steps and the step before
(time-centred).  All per-particle work runs in the HIP library; this module only
sequences calls and formats rows (the 8 CSV columns of antimatter-kernel.cc:128).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import (FP32, FP64, METHODS, SPECIES, Context, EnergyHistory, Store, _check, _p, lib, make_trap)

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


def simulate(params: dict, rng_seed: int = 42, precision: int = FP64, ctx: Context = None,
             coulomb_method: str = None, callback=None):
    """Run the reference driver loop on the GPU.  ``params`` uses the params.json keys.

    Returns (rows[k, 8], final state dict).  Columns: step, KE, PE_trap, PE_coul, total,
    n_active, fractional drift (of e*|U0|), absolute drift.
    """
    pops = params["particle_populations"]
    temps = params.get("particle_temperatures", {})
    rj = params.get("initial_cloud_radii", {})
    radii = (rj.get("radius_x", 1e-4), rj.get("radius_y", 1e-4), rj.get("radius_z", 1e-4))
    seed = int(params.get("seed", 42))
    b0 = float(params.get("b0", 1.0))
    u0 = float(params.get("trap_voltage", 10.0))
    field = make_trap(b0, u0, float(params.get("z0", 0.01)), float(params.get("rho0", 0.01)))
    steps = int(params.get("num_steps", 5000))
    cooling_scale = float(params.get("cooling_scale", 10.0))
    bath = float(params.get("bath_temperature", 0.5))
    coulomb_on = bool(params.get("coulomb_forces_enabled", True))
    method = coulomb_method or params.get("coulomb_computation_method", "standard")
    energy_method = params.get("energy_computation_method", "standard")
    cutoff = float(params.get("cutoff_radius", 0.0))
    bh_theta = float(params.get("barnes_hut_theta", 0.5))
    fmm_theta = float(params.get("fast_multipole_theta", 0.5))
    fmm_order = float(params.get("fast_multipole_order", 4))
    ann_on = bool(params.get("annihilation_enabled", False))
    ra = float(params.get("contact_radius", 1e-9))
    pa = float(params.get("annihilation_probability", 0.1))
    dt = derive_dt(pops, b0)

    pos, vel, species = initial_conditions(pops, temps, radii, seed)
    own = ctx is None
    ctx = ctx or Context(0)
    store = Store(ctx, len(species))
    store.upload(pos, vel, species)

    interval = steps // 100 if steps >= 100 else 1
    hist = EnergyHistory()

    def evaluate():
        e = store.compute_total_energy(field, energy_method, cutoff, bh_theta, fmm_theta, fmm_order, hist)
        pc = e.energy_potential if coulomb_on else 0.0
        return [e.kinetic, e.trap_potential, pc, e.num_active_particles]

    e0 = evaluate()
    total0 = e0[0] + e0[1] + e0[2]
    scale = E_CHARGE * abs(u0)
    if scale <= 0.0:
        scale = abs(total0)
    if scale <= 0.0:
        scale = 1.0
    rows = [[0, e0[0], e0[1], e0[2], total0, e0[3], 0.0, 0.0]]
    for step in range(1, steps + 1):
        if coulomb_on:
            if method == "barnes-hut":
                store.compute_barnes_hut_coulomb_forces(bh_theta, precision)
            elif method == "fast-multipole":
                store.compute_fast_multipole_coulomb_forces(int(fmm_order), fmm_theta)
            elif method == "all-pairs":
                store.compute_allpairs_coulomb_forces(precision)
            else:
                store.compute_standard_coulomb_forces(cutoff, precision)
        store.push(field, dt, cooling_scale, bath, rng_seed, step)
        if ann_on:
            store.annihilation(ra, pa, rng_seed, step)
        log = (step % interval == 0) or (step == steps)
        need = log or ((step + 1) % interval == 0) or (step + 1 == steps)
        if need:
            e = evaluate()
            if log:
                tot = e[0] + e[1] + e[2]
                absd = tot - total0
                rows.append([step, e[0], e[1], e[2], tot, e[3], absd / scale, absd])
        if callback is not None:
            callback(step, store)
    final = store.download()
    store.close()
    if own:
        ctx.close()
    return np.array(rows, dtype=np.float64), final
