"""FMM on the GPU against the oracle restatement of fast-multipole.cc."""
import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu


def cloud(n, seed=31, frac_neutral=0.05, frac_dead=0.04):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 3)) * np.array([1e-3, 1e-3, 4e-3])
    sp = rng.choice([0, 1, 3, 4], n).astype(np.uint8)
    sp[rng.random(n) < frac_neutral] = 2
    ann = (rng.random(n) < frac_dead).astype(np.uint8)
    return pos, sp, ann


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture
def store(ctx):
    s = am.Store(ctx, 20000)
    yield s
    s.close()


@pytest.mark.parametrize("order,theta", [(4, 0.5), (2, 0.7), (6, 0.3), (1, 0.5), (8, 0.5)])
def test_fmm_forces_match_oracle(store, order, theta):
    pos, sp, ann = cloud(5000)
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    store.compute_fast_multipole_coulomb_forces(order, theta)
    f = store.forces()
    m, q = oracle.mass_charge(sp)
    ref, st = oracle.fmm_forces(pos, q, ann, order, theta, want_stats=True)
    live = ann == 0
    gs = store.fmm_stats()
    assert gs["cells"] == st["cells"]
    assert gs["m2l_pairs"] == st["m2l_pairs"]
    np.testing.assert_allclose(f[live], ref[live], rtol=1e-9, atol=1e-10 * np.abs(ref).max())
    assert np.all(f[(sp == 2) & live] == 0)


def test_fmm_theta_zero_is_allpairs(store):
    pos, sp, ann = cloud(3000, seed=32, frac_neutral=0.0, frac_dead=0.0)
    store.upload(pos, np.zeros_like(pos), sp)
    store.compute_fast_multipole_coulomb_forces(4, 0.0)
    m, q = oracle.mass_charge(sp)
    assert rel_l2(store.forces(), oracle.allpairs_forces(pos, q)) < 1e-12
    assert store.fmm_stats()["m2l_pairs"] == 0


def test_fmm_potential_energy(store):
    pos, sp, ann = cloud(4000, seed=33)
    vel = np.zeros_like(pos)
    store.upload(pos, vel, sp, annihilated=ann)
    trap = am.make_trap(1.0, 10.0, 0.01, 0.01)
    m, q = oracle.mass_charge(sp)
    e = store.compute_total_energy(trap, "fast-multipole", fast_multipole_order=4, fast_multipole_theta=0.5)
    ref = oracle.fmm_potential_energy(pos, q, ann, 4, 0.5)
    assert e.energy_potential == pytest.approx(ref, rel=1e-10)


def test_fmm_small_and_degenerate(store):
    for n in (1, 2, 33):
        pos, sp, ann = cloud(n, seed=34 + n, frac_neutral=0.0, frac_dead=0.0)
        store.upload(pos, np.zeros_like(pos), sp)
        store.compute_fast_multipole_coulomb_forces(4, 0.5)
        m, q = oracle.mass_charge(sp)
        np.testing.assert_allclose(store.forces(), oracle.fmm_forces(pos, q, None, 4, 0.5), rtol=1e-10, atol=1e-30)
    store.upload(np.zeros((4, 3)), np.zeros((4, 3)), np.array([2, 2, 2, 2], np.uint8))  # all neutral
    store.compute_fast_multipole_coulomb_forces(4, 0.5)
    assert np.all(store.forces() == 0)


def test_fmm_fp32_near_field(store):
    pos, sp, ann = cloud(6000, seed=35, frac_neutral=0.0, frac_dead=0.0)
    store.upload(pos, np.zeros_like(pos), sp)
    m, q = oracle.mass_charge(sp)
    store.compute_fast_multipole_coulomb_forces(4, 0.5, am.FP32)
    f32 = store.forces()
    ref = oracle.fmm_forces(pos, q, None, 4, 0.5)
    assert rel_l2(f32, ref) < 1e-6  # FP32 near field and M2L arithmetic on top of the method
    store.compute_fast_multipole_coulomb_forces(4, 0.0, am.FP32)
    assert rel_l2(store.forces(), oracle.allpairs_forces(pos, q)) < 1e-5


@pytest.mark.parametrize("order,theta", [(1, 0.5), (2, 0.5), (3, 0.6), (4, 0.5), (5, 0.5), (6, 0.4), (7, 0.5),
                                         (8, 0.5)])
def test_fmm_fp32_orders_with_dead_and_neutral(store, order, theta):
    """The FP32 path for every order: the near field, and the M2L on matrix cores for orders <= 5
    (36-coefficient rows up to order 4, 56 at order 5; the FP64 contraction above), with annihilated
    and neutral particles (sources of charge 0 in the near lists) and leaves of every occupancy
    class."""
    pos, sp, ann = cloud(7000, seed=40 + order)
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    store.compute_fast_multipole_coulomb_forces(order, theta, am.FP32)
    f = store.forces()
    m, q = oracle.mass_charge(sp)
    ref = oracle.fmm_forces(pos, q, ann, order, theta)
    live = ann == 0
    assert rel_l2(f[live], ref[live]) < 1e-6
    assert np.all(f[(sp == 2) & live] == 0)
    assert np.all(f[~live] == 0)


def test_fmm_fp32_overfull_leaf(store):
    """45 coincident particles fill one deepest-level leaf past 32: the FP32 near field takes its
    targets in batches of 32 (coincident pairs contribute nothing, as in the reference)."""
    pos, sp, ann = cloud(3000, seed=47, frac_neutral=0.0, frac_dead=0.0)
    pos = np.concatenate([pos, np.tile(np.array([[2e-4, -1e-4, 5e-4]]), (45, 1))])
    sp = np.concatenate([sp, np.where(np.arange(45) % 2 == 0, 1, 0).astype(np.uint8)])
    store.upload(pos, np.zeros_like(pos), sp)
    m, q = oracle.mass_charge(sp)
    store.compute_fast_multipole_coulomb_forces(4, 0.5, am.FP32)
    ref = oracle.fmm_forces(pos, q, None, 4, 0.5)
    assert rel_l2(store.forces(), ref) < 1e-6
    np.testing.assert_allclose(store.forces()[-45:], ref[-45:], rtol=1e-5, atol=1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("order,theta", [(4, 0.5), (5, 0.4), (3, 0.6)])
def test_fmm_fp32_m2l_paths_agree(ctx, order, theta):
    """The three M2L routes of the FP32 path on one cloud: displacement classes on matrix cores
    (the default), every pair with a tensor of its own on matrix cores (classes switched off, so
    every pair takes the unclassed route), and the FP64 contraction (the route past the unclassed
    bound).  Each is within the FP32 bound of the oracle; the two matrix-core routes form the same
    FP32 products and differ only by their exact-accumulation grids (the grid follows a group's
    product count, 16 times larger when every pair is its own product); repeated calls are
    bit-identical (the accumulation is exact, so its order does not matter)."""
    pos, sp, ann = cloud(7000, seed=60 + order)
    s = am.Store(ctx, len(sp))
    s.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    m, q = oracle.mass_charge(sp)
    ref = oracle.fmm_forces(pos, q, ann, order, theta)
    live = ann == 0
    out = {}
    try:
        for name, classes, cap in (("classes", True, 1 << 20), ("unclassed", False, 1 << 20), ("fp64", True, -1)):
            ctx.fmm_m2l_options(classes, cap)
            s.compute_fast_multipole_coulomb_forces(order, theta, am.FP32)
            out[name] = (s.forces(), s.fmm_stats())
            print(name, out[name][1], flush=True)
            assert rel_l2(out[name][0][live], ref[live]) < 1e-6, name
    finally:
        ctx.fmm_m2l_options(True, 1 << 20)
    st = out["classes"][1]
    assert st["m2l_pairs"] > 0 and st["m2l_unclassed"] == 0 and st["m2l_items"] >= 1
    assert out["unclassed"][1]["m2l_unclassed"] == st["m2l_pairs"] and out["unclassed"][1]["m2l_items"] >= 1
    assert out["fp64"][1]["m2l_items"] == 0
    assert rel_l2(out["unclassed"][0][live], out["classes"][0][live]) < 1e-10
    s.compute_fast_multipole_coulomb_forces(order, theta, am.FP32)
    np.testing.assert_array_equal(s.forces(), out["classes"][0])
    s.close()
