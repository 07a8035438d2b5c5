"""Barnes-Hut on the GPU against the oracle restatement of barnes-hut.cc."""
import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu


def cloud(n, seed=21, frac_neutral=0.05, frac_dead=0.04):
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


def test_tree_order_and_leaves_match_reference(store):
    pos, sp, ann = cloud(6000)
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    store.compute_barnes_hut_coulomb_forces(0.5, am.FP64)
    t = store.bh_tree()
    alive = np.where(ann == 0)[0]
    m, q = oracle.mass_charge(sp)
    ref = oracle.bh_tree(pos[alive], q[alive], 8)
    # gathered index -> particle index
    np.testing.assert_array_equal(t["order"], alive[ref["order"]])
    np.testing.assert_array_equal(t["begin"], ref["begin"])
    np.testing.assert_array_equal(t["end"], ref["end"])
    np.testing.assert_array_equal(t["leaf"], ref["leaf"])
    np.testing.assert_array_equal(t["qp"], ref["qp"])
    np.testing.assert_array_equal(t["qn"], ref["qn"])
    assert t["deep_overflow"] == 0


def test_tree_depth_changes_between_builds(store):
    # Shallow and deep trees built one after another on one context.  A build first launches its
    # levels to one below the previous tree's depth and must carry on when the new tree is deeper.
    rng = np.random.default_rng(31)
    base = rng.uniform(-1, 1, (3000, 3)) * 1e-3
    deep = base.copy()
    deep[:150] = 2e-4 + rng.uniform(-1, 1, (150, 3)) * 1e-8  # a tight clump: about 16 levels deeper
    sp = rng.choice([0, 1], 3000).astype(np.uint8)
    m, q = oracle.mass_charge(sp)
    depths = []
    for pos in (base, deep, base, deep, deep):
        store.upload(pos, np.zeros_like(pos), sp)
        store.compute_barnes_hut_coulomb_forces(0.5, am.FP64)
        t = store.bh_tree()
        ref = oracle.bh_tree(pos, q, 8)
        np.testing.assert_array_equal(t["order"], ref["order"])
        np.testing.assert_array_equal(t["begin"], ref["begin"])
        np.testing.assert_array_equal(t["end"], ref["end"])
        np.testing.assert_array_equal(t["leaf"], ref["leaf"])
        np.testing.assert_array_equal(t["qp"], ref["qp"])
        np.testing.assert_array_equal(t["qn"], ref["qn"])
        assert t["deep_overflow"] == 0
        depths.append(int(ref["depth"].max()))
    assert depths[1] > depths[0] and depths[2] == depths[0] and depths[3] == depths[1]


@pytest.mark.parametrize("theta", [0.0, 0.5, 0.9])
def test_fp64_forces_and_counts_bit_identical(store, theta):
    pos, sp, ann = cloud(4000, seed=22)
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    cnt = store.barnes_hut_forces_counted(theta, am.FP64)
    f = store.forces()
    m, q = oracle.mass_charge(sp)
    ref, rcnt = oracle.bh_forces(pos, q, ann, theta, 8, want_counts=True)
    live = ann == 0
    np.testing.assert_array_equal(cnt[live], rcnt[live])
    np.testing.assert_array_equal(f[live], ref[live])


def lattice_cloud(shift=(0.0, 0.0, 0.0)):
    # 17^3 particles on a dyadic lattice (spacing 2^-14 m, centred on the origin), alternating e+/e-.
    # The root cube is the bounding box widened by (1 + 1e-9), so cube centres and widths are lattice
    # points up to that factor: 500-1200 (target, internal cube) pairs per theta lie within 1e-8
    # relative of the acceptance threshold, inside the FP32 test's undecided band of 2^-16, where
    # the FP64 test decides.
    g = np.arange(-8, 9) * 2.0 ** -14
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.asarray(shift)
    sp = np.where(np.arange(len(pos)) % 2 == 0, 1, 3).astype(np.uint8)
    return pos, sp


@pytest.mark.parametrize("theta", [0.25, 0.5, 1.0])
def test_acceptance_near_ties_on_a_lattice(store, theta):
    pos, sp = lattice_cloud()
    store.upload(pos, np.zeros_like(pos), sp)
    m, q = oracle.mass_charge(sp)
    ref, rcnt = oracle.bh_forces(pos, q, None, theta, 8, want_counts=True)
    cnt64 = store.barnes_hut_forces_counted(theta, am.FP64)
    np.testing.assert_array_equal(cnt64, rcnt)
    np.testing.assert_array_equal(store.forces(), ref)
    cnt32 = store.barnes_hut_forces_counted(theta, am.FP32)
    np.testing.assert_array_equal(cnt32, rcnt)
    assert rel_l2(store.forces(), ref) < 1e-5


def test_cloud_far_from_the_origin_takes_the_fp64_test(store):
    # 1 m away from the origin the FP32 test's range check fails (the coordinates exceed 8 root
    # half widths), so every acceptance test is the FP64 one; the interaction sets stay the oracle's.
    pos, sp, ann = cloud(4000, seed=24)
    pos = pos + np.array([1.0, -0.5, 0.25])
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    m, q = oracle.mass_charge(sp)
    ref, rcnt = oracle.bh_forces(pos, q, ann, 0.6, 8, want_counts=True)
    live = ann == 0
    cnt = store.barnes_hut_forces_counted(0.6, am.FP64)
    np.testing.assert_array_equal(cnt[live], rcnt[live])
    np.testing.assert_array_equal(store.forces()[live], ref[live])
    cnt = store.barnes_hut_forces_counted(0.6, am.FP32)
    np.testing.assert_array_equal(cnt[live], rcnt[live])


def test_fp32_theta0_is_allpairs_and_theta_error(store):
    pos, sp, ann = cloud(8000, seed=23, frac_neutral=0.0, frac_dead=0.0)
    store.upload(pos, np.zeros_like(pos), sp)
    m, q = oracle.mass_charge(sp)
    ap = oracle.allpairs_forces(pos, q)
    store.compute_barnes_hut_coulomb_forces(0.0, am.FP32)
    assert rel_l2(store.forces(), ap) < 1e-5
    store.compute_barnes_hut_coulomb_forces(0.7, am.FP32)
    f32 = store.forces()
    ref = oracle.bh_forces(pos, q, None, 0.7)
    assert rel_l2(f32, ref) < 1e-5  # FP32 arithmetic error on top of the method error
    err = rel_l2(ref, ap)
    print(f"BH theta=0.7 method error vs all-pairs: {err:.3e}")
    assert 1e-3 < err < 5e-2


def test_potential_energy_matches_oracle(store):
    pos, sp, ann = cloud(3000, seed=24)
    vel = np.random.default_rng(1).normal(0, 1e3, pos.shape)
    store.upload(pos, vel, sp, annihilated=ann)
    trap = am.make_trap(1.0, 10.0, 0.01, 0.01)
    m, q = oracle.mass_charge(sp)
    for theta in (0.0, 0.6):
        e = store.compute_total_energy(trap, "barnes-hut", barnes_hut_theta=theta)
        ref = oracle.bh_potential_energy(pos, q, ann, theta)
        assert e.energy_potential == pytest.approx(ref, rel=1e-12)


def test_degenerate_inputs(store):
    # one particle, all particles at one point, all annihilated
    store.upload(np.zeros((1, 3)), np.zeros((1, 3)), np.array([1], np.uint8))
    store.compute_barnes_hut_coulomb_forces(0.5, am.FP32)
    assert np.all(store.forces() == 0)
    pos = np.full((20, 3), 1e-4)
    sp = np.array([1, 3] * 10, np.uint8)
    store.upload(pos, np.zeros_like(pos), sp)
    store.compute_barnes_hut_coulomb_forces(0.5, am.FP64)
    assert np.all(store.forces() == 0)  # coincident pairs are skipped (r^2 < 1e-30)
    store.upload(pos, np.zeros_like(pos), sp, annihilated=np.ones(20, np.uint8))
    store.compute_barnes_hut_coulomb_forces(0.5, am.FP32)


@pytest.mark.parametrize("species", ["mixed", "positrons"])
def test_fp32_neutral_dead_and_one_sign(store, species):
    """FP32 walk with neutral and annihilated particles (charge-0 leaf sources, nodes whose
    negative class is empty), against the FP64 oracle; the interaction counts of the FP32 walk
    equal the FP64 walk's (same MAC decisions, charge-0 sources and the target itself excluded)."""
    pos, sp, ann = cloud(6000, seed=25)
    if species == "positrons":
        sp = np.where(sp == 2, 2, 1).astype(np.uint8)  # positrons and neutrals only
    store.upload(pos, np.zeros_like(pos), sp, annihilated=ann)
    m, q = oracle.mass_charge(sp)
    cnt32 = store.barnes_hut_forces_counted(0.6, am.FP32)
    f32 = store.forces()
    cnt64 = store.barnes_hut_forces_counted(0.6, am.FP64)
    ref = oracle.bh_forces(pos, q, ann, 0.6, 8)
    live = ann == 0
    assert rel_l2(f32[live], ref[live]) < 1e-5
    assert np.all(f32[(sp == 2) & live] == 0)
    np.testing.assert_array_equal(cnt32[live], cnt64[live])
