"""GPU parity of the direct solvers, push, annihilation and energy against the CPU oracle.

Tolerances: FP64 paths are compared bit for bit where the summation order is the oracle's
(all-pairs, push without cooling) and to 1e-13 relative otherwise; the FP32 all-pairs kernel to
the 1e-5 global relative force error of the all-pairs workload's accuracy gate."""
import json
import os

import numpy as np
import pytest

import antimatter_amd as am
import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

E = oracle.E_CHARGE


def config2_cloud(n, seed=42):
    """The all-pairs workload's cloud: half antiprotons (10 K), half positrons (100 K)."""
    pops = {"antiproton": n // 2, "positron": n - n // 2}
    temps = {"antiproton": 10.0, "positron": 100.0}
    return am.initial_conditions(pops, temps, (1e-3, 1e-3, 4e-3), seed)


def mixed_cloud(n, seed=3, radius=2e-5):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 3)) * radius
    vel = rng.normal(0, 1e4, (n, 3))
    sp = rng.choice([0, 1, 2, 3, 4], n).astype(np.uint8)
    return pos, vel, sp


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def upload(ctx, pos, vel, sp, ann=None, cap=None):
    s = am.Store(ctx, cap or max(1, len(sp)))
    s.upload(pos, vel, sp, annihilated=ann)
    return s


def test_allpairs_fp64_bit_identical_to_oracle(ctx):
    for n in (1, 2, 257, 1000):
        pos, vel, sp = mixed_cloud(n)
        ann = np.zeros(n, np.uint8)
        ann[::7] = 1
        s = upload(ctx, pos, vel, sp, ann)
        s.compute_allpairs_coulomb_forces(am.FP64)
        f = s.forces()
        m, q = oracle.mass_charge(sp)
        ref = oracle.allpairs_forces(pos, q, ann)
        live = ann == 0
        np.testing.assert_array_equal(f[live], ref[live])
        s.close()


def test_allpairs_fp32_meets_1e5_on_config2_cloud(ctx):
    n = 8192
    pos, vel, sp = config2_cloud(n)
    s = upload(ctx, pos, vel, sp)
    s.compute_allpairs_coulomb_forces(am.FP32)
    f = s.forces()
    m, q = oracle.mass_charge(sp)
    ref = oracle.allpairs_forces(pos, q)
    err = rel_l2(f, ref)
    per = np.linalg.norm(f - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"N={n} FP32 all-pairs: global {err:.3e} median {np.median(per):.3e} max {per.max():.3e}")
    assert err < 1e-5
    s.close()


def test_allpairs_fp32_full_config2_size(ctx):
    """N = 65,536 (the all-pairs workload): FP32 vs the GPU FP64 kernel over all particles, and
    vs the oracle on a seeded subset of targets."""
    n = 65536
    pos, vel, sp = config2_cloud(n)
    s = upload(ctx, pos, vel, sp)
    s.compute_allpairs_coulomb_forces(am.FP32)
    f32 = s.forces()
    s.compute_allpairs_coulomb_forces(am.FP64)
    f64 = s.forces()
    m, q = oracle.mass_charge(sp)
    tgt = np.random.default_rng(0).choice(n, 256, replace=False)
    ref = oracle.allpairs_forces_subset(pos, q, None, tgt)
    np.testing.assert_allclose(f64[tgt], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    err = rel_l2(f32, f64)
    print(f"N={n} FP32 vs FP64 global rel L2 {err:.3e}")
    assert err < 1e-5
    # Newton's third law on the whole cloud (size-independent property).
    assert np.linalg.norm(f32.sum(0)) < 1e-6 * np.abs(f32).sum(0).max()
    s.close()


def test_allpairs_fp32_sorted_path_with_dead_and_neutral(ctx):
    """The spatially sorted kernel (N >= 4096): annihilated and neutral particles, repeated calls
    (cached order), then a re-upload of a different cloud."""
    rng = np.random.default_rng(40)
    for n in (5000, 12345):
        pos = rng.uniform(-1, 1, (n, 3)) * np.array([1e-3, 2e-3, 4e-3])
        sp = rng.choice([0, 1, 2, 3, 4], n).astype(np.uint8)
        ann = (rng.random(n) < 0.1).astype(np.uint8)
        s = upload(ctx, pos, np.zeros_like(pos), sp, ann, cap=20000)
        m, q = oracle.mass_charge(sp)
        ref = oracle.allpairs_forces(pos, q, ann)
        for _ in range(3):
            s.compute_allpairs_coulomb_forces(am.FP32)
            f = s.forces()
            live = ann == 0
            assert rel_l2(f[live], ref[live]) < 1e-5
            assert np.all(f[(sp == 2) & live] == 0)
        s.close()


def test_allpairs_sorted_path_non_unit_charges(ctx):
    """Tiles holding charges other than +-e (antihelium, caller overrides) take the generic pair
    loop; unit-charge tiles fold the sign into the difference.  Both against the oracle, with the
    target order of the tiles partitioned by sign."""
    rng = np.random.default_rng(41)
    n = 9000
    pos = rng.normal(0, 1, (n, 3)) * np.array([1e-3, 1e-3, 3e-3])
    sp = rng.choice([0, 1, 3, 4], n).astype(np.uint8)
    sp[:300] = 5  # antihelium (-2e) in the first tiles of some regions
    m, q = oracle.mass_charge(sp)
    s = upload(ctx, pos, np.zeros_like(pos), sp)
    s.compute_allpairs_coulomb_forces(am.FP32)
    ref = oracle.allpairs_forces(pos, q)
    assert rel_l2(s.forces(), ref) < 1e-5
    q2 = q.copy()
    q2[::17] *= 0.37  # fractional charges via the upload override
    s.upload(pos, np.zeros_like(pos), sp, charge=q2)
    for _ in range(2):
        s.compute_allpairs_coulomb_forces(am.FP32)
        ref2 = oracle.allpairs_forces(pos, q2)
        assert rel_l2(s.forces(), ref2) < 1e-5
    s.close()


def test_allpairs_edge_cases(ctx):
    s = am.Store(ctx, 16)
    s.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint8))
    s.compute_allpairs_coulomb_forces(am.FP32)  # empty store: no-op
    pos, vel, sp = mixed_cloud(5)
    s.upload(pos, vel, sp, annihilated=np.ones(5, np.uint8))
    s.compute_allpairs_coulomb_forces(am.FP32)
    assert np.all(s.forces() == 0)
    # neutral particles (antihydrogen) feel and exert no force
    sp = np.array([1, 2, 3], np.uint8)
    pos = np.array([[0, 0, 0], [1e-6, 0, 0], [0, 2e-6, 0]], dtype=np.float64)
    s.upload(pos, np.zeros((3, 3)), sp)
    s.compute_allpairs_coulomb_forces(am.FP32)
    f = s.forces()
    assert np.all(f[1] == 0)
    ref = oracle.allpairs_forces(pos, oracle.mass_charge(sp)[1])
    np.testing.assert_allclose(f, ref, rtol=1e-6)
    s.close()


def test_push_matches_oracle(ctx):
    n = 1000
    pos, vel, sp = mixed_cloud(n, radius=1e-3)
    m, q = oracle.mass_charge(sp)
    ann = np.zeros(n, np.uint8)
    ann[::9] = 1
    force = np.random.default_rng(1).normal(0, 1e-18, (n, 3))
    trap = am.make_trap(1.0, 10.0, 0.01, 0.01)
    dt = 8.930967e-12
    s = upload(ctx, pos, vel, sp, ann)
    s.set_coulomb_force(force)
    s.push(trap, dt)  # Lorentz only (tau_scale 0)
    s.set_coulomb_force(force)
    s.push(trap, dt)
    g = s.download()
    p2, v2, f2 = pos.copy(), vel.copy(), force.copy()
    st = np.zeros(n, np.uint8)
    for _ in range(2):
        f2[:] = force
        oracle.lorentz_push(p2, v2, f2, q, m, ann, st, 1.0, 10.0, trap.d, dt)
    np.testing.assert_array_equal(g["pos"], p2)
    np.testing.assert_array_equal(g["vel"], v2)
    np.testing.assert_array_equal(g["staggered"][ann == 0], 1)
    assert np.all(g["force"][ann == 0] == 0)
    # with cooling (scale 1e-9 to make it visible) and the bath noise
    s.push(trap, dt, 1e-9, 4.0, 7, 3)
    g = s.download()
    f2[:] = 0
    oracle.lorentz_push(p2, v2, f2, q, m, ann, st, 1.0, 10.0, trap.d, dt)
    oracle.cyclotron_cooling(v2, q, m, ann, 1.0, 1e-9, 4.0, dt, 7, 3)
    np.testing.assert_allclose(g["vel"], v2, rtol=1e-13, atol=1e-12)
    np.testing.assert_array_equal(g["pos"], p2)
    s.close()


def test_standard_forces_match_oracle(ctx):
    params = json.load(open(os.path.join(GOLDEN, "params_config1.json")))
    pos, vel, sp = am.initial_conditions(params["particle_populations"], params["particle_temperatures"],
                                         (1e-4, 1e-4, 1e-4), 42)
    m, q = oracle.mass_charge(sp)
    s = upload(ctx, pos, vel, sp)
    for rc in (1e-5, 3e-5, 1.0):
        s.compute_standard_coulomb_forces(rc)
        f = s.forces()
        ref = oracle.standard_forces(pos, q, None, rc)
        np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13 * max(np.abs(ref).max(), 1e-300))
    s.compute_standard_coulomb_forces(0.0)
    assert np.all(s.forces() == 0)
    s.close()


@pytest.mark.parametrize("p", [1.0, 0.5])
def test_annihilation_matches_oracle(ctx, p):
    rng = np.random.default_rng(12)
    n = 3000
    pos = rng.uniform(-1, 1, (n, 3)) * 1e-4
    sp = rng.choice([0, 1, 2, 3, 4], n).astype(np.uint8)
    ann = np.zeros(n, np.uint8)
    ann[::13] = 1
    s = upload(ctx, pos, np.zeros((n, 3)), sp, ann)
    k = s.annihilation(6e-6, p, 99, 5)
    g = s.download()
    a2 = ann.copy()
    cnt, _ = oracle.annihilate(pos, sp, a2, 6e-6, p, 99, 5)
    assert k == cnt and cnt > 0
    np.testing.assert_array_equal(g["annihilated"], a2)
    s.close()


@pytest.mark.parametrize("case", ["clumps", "chain", "long-list"])
@pytest.mark.parametrize("p", [1.0, 0.5])
def test_annihilation_dense_cases_match_oracle(ctx, case, p):
    # clumps: several candidates per particle, so the greedy choices interact; chain: a line of
    # alternating e+/e- at 0.9 contact radii, numbered along the line, so each decision waits for
    # the previous one all the way down (longer than the parallel rounds go); long-list: more
    # candidate entries than the in-LDS resolution holds.
    rng = np.random.default_rng(40)
    ra = 6e-6
    if case == "clumps":
        n = 2500  # about 4,500 candidate entries, up to 8 per particle
        centres = rng.uniform(-1, 1, (150, 3)) * 1e-4
        pos = centres[rng.integers(0, 150, n)] + rng.normal(size=(n, 3)) * 3e-6
        sp = rng.choice([0, 1, 3, 4], n).astype(np.uint8)
    elif case == "chain":
        n = 300
        pos = np.zeros((n, 3))
        pos[:, 0] = np.arange(n) * 0.9 * ra
        sp = np.where(np.arange(n) % 2 == 0, 1, 3).astype(np.uint8)
    else:
        n = 20000
        pos = rng.uniform(-1, 1, (n, 3)) * 4e-5
        sp = rng.choice([1, 3], n).astype(np.uint8)
    ann = np.zeros(n, np.uint8)
    s = upload(ctx, pos, np.zeros((n, 3)), sp, ann)
    k = s.annihilation(ra, p, 7, 3)
    g = s.download()
    a2 = ann.copy()
    cnt, _ = oracle.annihilate(pos, sp, a2, ra, p, 7, 3)
    assert k == cnt and cnt > 0
    np.testing.assert_array_equal(g["annihilated"], a2)
    s.close()


def test_energy_matches_oracle(ctx):
    n = 700
    pos, vel, sp = mixed_cloud(n, radius=3e-5)
    m, q = oracle.mass_charge(sp)
    ann = np.zeros(n, np.uint8)
    ann[::5] = 1
    trap = am.make_trap(1.0, 10.0, 0.01, 0.01)
    s = upload(ctx, pos, vel, sp, ann)
    for method, kw in (("standard", dict(rc=1e-5)), ("all-pairs", {})):
        e = s.compute_total_energy(trap, method, cutoff_radius=kw.get("rc", 0.0))
        ref = oracle.energy(pos, vel, q, m, ann, 10.0, trap.d, method=0 if method == "standard" else 1,
                            rc=kw.get("rc", 0.0), bh_theta=0.0)
        assert e.num_active_particles == ref["num_active_particles"]
        assert e.kinetic == pytest.approx(ref["kinetic"], rel=1e-13)
        assert e.trap_potential == pytest.approx(ref["trap_potential"], rel=1e-12)
        assert e.energy_potential == pytest.approx(ref["energy_potential"], rel=1e-12)
    s.close()


def test_config1_driver_matches_notebook_and_oracle(ctx):
    params = json.load(open(os.path.join(GOLDEN, "params_config1.json")))
    golden = json.load(open(os.path.join(GOLDEN, "config1_notebook.json")))
    rows = am.simulate(params)
    orows, ofin = oracle.simulate(params)
    assert rows.shape == orows.shape == (101, 8)
    np.testing.assert_array_equal(rows[:, 5], orows[:, 5])  # identical annihilation history
    np.testing.assert_allclose(rows[:, 1:5], orows[:, 1:5], rtol=1e-9)
    for g in golden["rows"]:
        r = rows[g["step"]]
        for k, col in (("kinetic", 1), ("trap_potential", 2), ("energy_potential", 3), ("total", 4)):
            assert float("%.5e" % r[col]) == pytest.approx(g[k], rel=1.01e-5)
    drift = rows[:, 6]
    windows = [drift[k + 20] - drift[k] for k in (0, 20, 40, 60, 80)]
    np.testing.assert_allclose(windows, golden["drift_change_per_20_steps"], atol=1.5e-6)
    assert int(rows[-1, 5]) == golden["active_end"]
