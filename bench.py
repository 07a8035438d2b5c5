#!/usr/bin/env python3
"""Benchmark of the all-pairs hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): direct all-pairs Coulomb without
cutoff, N = 65,536 (32,768 antiprotons at 10 K + 32,768 positrons at 100 K, ellipsoidal cloud
radii (1e-3, 1e-3, 4e-3) m), B0 = 3 T, U0 = 50 V, z0 = rho0 = 0.01 m, dt from the reference
rule.  One step = FP32 all-pairs force (hi/lo coordinates, FP64 tile accumulation) + fused FP64
Boris push, with every input resident in HBM.  Energy diagnostics and the force-error check run
outside the timed region.

Prints one JSON line (rank 0).  --gpus N > 1 exits: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "antimatter-kernel_amd"))
sys.path.insert(0, ROOT)

N_PARTICLES = 65536
FLOP_PER_PAIR = 20  # Nyland-Harris-Prins convention: one evaluated ordered pair interaction
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak
HBM_PEAK_GBS = 8000.0


def workload(n=N_PARTICLES, seed=42):
    import antimatter_amd as am
    pops = {"antiproton": n // 2, "positron": n - n // 2}
    temps = {"antiproton": 10.0, "positron": 100.0}
    pos, vel, sp = am.initial_conditions(pops, temps, (1e-3, 1e-3, 4e-3), seed)
    trap = am.make_trap(3.0, 50.0, 0.01, 0.01)
    dt = am.derive_dt(pops, 3.0)
    return pos, vel, sp, trap, dt


def cpu_baseline(pos, vel, sp, trap, dt, steps=4):
    """The oracle's OpenMP FP64 all-pairs + Boris (same workload), median per-step time."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    oracle.set_num_threads(threads)
    m, q = oracle.mass_charge(sp)
    p, v = pos.copy(), vel.copy()
    f = np.zeros_like(p)
    ann = np.zeros(len(sp), np.uint8)
    stag = np.zeros(len(sp), np.uint8)
    times = []
    for k in range(steps + 1):  # one warm-up step
        t0 = time.perf_counter()
        f[:] = oracle.allpairs_forces(p, q, ann)
        oracle.lorentz_push(p, v, f, q, m, ann, stag, trap.b0, trap.trap_voltage, trap.d, dt)
        t1 = time.perf_counter()
        if k > 0:
            times.append(t1 - t0)
    med = statistics.median(times)
    return dict(value=len(sp) / med, unit="particle-steps/s", cores=oracle.max_threads(), kind="port",
                sample=f"{steps} steps (after 1 warm-up) of the same N={len(sp)} all-pairs FP64 force + Boris "
                       f"push, median {med:.3f} s/step, OpenMP")


def read_traffic():
    """HBM bytes per all-pairs launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", "allpairs_f32_n65536_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        return float(json.load(open(path))["hbm_bytes_per_launch"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the all-pairs path would shard targets across GPUs with an all-gather of positions).",
              file=sys.stderr)
        sys.exit(2)

    import antimatter_amd as am
    pos, vel, sp, trap, dt = workload()
    n = len(sp)
    ctx = am.Context(0)
    store = am.Store(ctx, n)
    store.upload(pos, vel, sp)
    hist = am.EnergyHistory()
    e0 = store.compute_total_energy(trap, "all-pairs", history=hist)

    def step(k):
        store.compute_allpairs_coulomb_forces(am.FP32)
        store.push(trap, dt, 0.0, 0.0, 42, k)

    for k in range(args.warmup):
        step(k + 1)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.enable_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k + 1)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    ms_force, launches = ctx.kernel_time("allpairs_f32")
    ms_push, _ = ctx.kernel_time("push")
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n * args.steps / elapsed

    # Accuracy at the final state: FP32 forces vs the FP64 all-pairs kernel (global L2).
    store.compute_allpairs_coulomb_forces(am.FP32)
    f32 = store.forces()
    store.compute_allpairs_coulomb_forces(am.FP64)
    f64 = store.forces()
    err = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
    per = np.linalg.norm(f32 - f64, axis=1) / np.maximum(np.linalg.norm(f64, axis=1), 1e-300)
    # Energy drift over the run (the reference's definition, antimatter-kernel.cc:164-177): the
    # potentials of a logged step are averaged with the step before it, so evaluate after the
    # last timed step and after one further (untimed) step.
    hist_end = am.EnergyHistory()
    store.compute_total_energy(trap, "all-pairs", history=hist_end)
    step(args.warmup + args.steps + 1)
    e1 = store.compute_total_energy(trap, "all-pairs", history=hist_end)
    drift = (e1.total() - e0.total()) / (1.602176634e-19 * abs(trap.trap_voltage))

    avg_ms = ms_force / max(launches, 1)
    flops = FLOP_PER_PAIR * n * (n - 1)
    achieved = flops / (avg_ms * 1e-3) / 1e12
    traffic = read_traffic()
    out = {
        "metric": "particle-steps/s at <=1e-5 rel force err vs FP64 direct; % FP32 peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (reference initial-condition generators, seed 42)",
        "config": {
            "workload": "direct all-pairs Coulomb (no cutoff) + FP64 Boris push, N=65536 "
                        "(32768 antiprotons 10 K + 32768 positrons 100 K), radii (1e-3,1e-3,4e-3) m, "
                        "B0=3 T, U0=50 V, z0=rho0=0.01 m",
            "n_particles": n,
            "dt_s": dt,
            "pair_arithmetic": "fp32: hi/lo coordinates for near tiles, tile-local coordinates for far tiles; "
                               "fp64 per-128-source tile accumulation",
            "update_arithmetic": "fp64",
        },
        "force_rel_err_l2": err,
        "force_rel_err_max": float(per.max()),
        "force_rel_err_median": float(np.median(per)),
        "energy_drift_over_run": drift,
        "pair_interactions_per_s": n * (n - 1) * args.steps / elapsed,
        "kernel_ms": {"allpairs_f32_avg": avg_ms, "push_avg": ms_push / max(launches, 1)},
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, vel, sp, trap, dt, steps=args.cpu_steps)
    store.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
