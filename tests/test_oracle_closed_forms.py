"""Pins of the CPU oracle by closed forms, invariants and brute force (each chosen so that a
plausible mistake — dropped term, wrong sign or index, transposed operand — fails)."""
import math

import numpy as np
import pytest

import oracle

K = oracle.K_COULOMB
E = oracle.E_CHARGE


def cloud(n, seed=1, radii=(1e-3, 1e-3, 4e-3), frac_pos=0.5):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 3)) * np.asarray(radii)
    q = np.where(rng.random(n) < frac_pos, E, -E)
    return pos, q


# ---- RNG --------------------------------------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10.
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in kat:
        assert tuple(int(x) for x in oracle.philox4x32_10(c, k)) == want


def test_cooling_normals_statistics():
    g = np.array([oracle.cooling_normals(7, 3, i) for i in range(20000)])
    assert abs(g.mean()) < 0.02
    assert g.std() == pytest.approx(1.0, abs=0.02)


# ---- field / Boris / cooling -------------------------------------------------------------------
def test_field_is_minus_grad_potential():
    d = oracle.characteristic_dimension(0.01, 0.01)
    assert d * d == pytest.approx(0.01 ** 2 / 2 + 0.01 ** 2 / 4, rel=1e-15)
    x = np.array([[1e-3, -2e-3, 3e-3]])
    e = oracle.electric_field(10.0, d, x)[0]
    h = 1e-7
    grad = []
    for k in range(3):
        xp, xm = x.copy(), x.copy()
        xp[0, k] += h
        xm[0, k] -= h
        grad.append((oracle.potential(10.0, d, xp)[0] - oracle.potential(10.0, d, xm)[0]) / (2 * h))
    np.testing.assert_allclose(e, -np.array(grad), rtol=1e-7)


def _push_once(v0, b0=1.0, u0=0.0, dt=1e-12, species=3, steps=1, staggered=1, x0=(0, 0, 0)):
    m, q = oracle.species_props(species)
    pos = np.array([x0], dtype=np.float64)
    vel = np.array([v0], dtype=np.float64)
    f = np.zeros((1, 3))
    qa, ma = np.array([q]), np.array([m])
    ann = np.zeros(1, np.uint8)
    st = np.array([staggered], np.uint8)
    d = oracle.characteristic_dimension(0.01, 0.01)
    for _ in range(steps):
        oracle.lorentz_push(pos, vel, f, qa, ma, ann, st, b0, u0, d, dt)
    return pos[0], vel[0], st[0]


def test_boris_pure_b_conserves_speed_and_rotates_by_2atan():
    m, q = oracle.species_props(3)
    wc = abs(q) * 1.0 / m
    dt = math.pi / (2 * wc)  # a quarter cyclotron period, the reference's dt rule
    v0 = np.array([1e5, 0.0, 2e4])
    _, v1, _ = _push_once(v0, dt=dt)
    assert np.linalg.norm(v1) == pytest.approx(np.linalg.norm(v0), rel=1e-14)
    ang = math.atan2(v1[1], v1[0])
    # Boris rotation angle 2 atan(w dt / 2) (76.3 degrees at a quarter period); electrons
    # (q < 0) rotate counter-clockwise about +z.
    assert abs(ang) == pytest.approx(2 * math.atan(wc * dt / 2), rel=1e-12)
    assert ang > 0
    assert v1[2] == pytest.approx(v0[2], rel=1e-15)


def test_boris_back_stagger_on_first_call_only():
    v0 = np.array([1e5, 0.0, 0.0])
    _, v_a, st = _push_once(v0, staggered=0)
    assert st == 1
    _, v_b, _ = _push_once(v0, staggered=1)
    assert not np.allclose(v_a, v_b)


def test_boris_axial_harmonic_frequency():
    """Positrons at U0 > 0 oscillate axially with w_z^2 = q U0 / (m d^2) (leapfrog: 2 asin(w dt/2))."""
    d = oracle.characteristic_dimension(0.01, 0.01)
    m, q = oracle.species_props(1)
    wz = math.sqrt(q * 10.0 / (m * d * d))
    dt = 0.05 / wz
    pos = np.array([[0.0, 0.0, 1e-4]])
    vel = np.zeros((1, 3))
    f = np.zeros((1, 3))
    st = np.array([0], np.uint8)
    zs = []
    for _ in range(400):
        oracle.lorentz_push(pos, vel, f, np.array([q]), np.array([m]), np.zeros(1, np.uint8), st, 1.0, 10.0, d, dt)
        zs.append(pos[0, 2])
    zs = np.array(zs)
    cross = np.where(np.diff(np.sign(zs)) != 0)[0]
    period_steps = 2 * np.mean(np.diff(cross))
    w_num = 2 * math.pi / (period_steps * dt)
    assert w_num == pytest.approx(2 * math.asin(wz * dt / 2) / dt, rel=2e-3)


def test_cooling_time_known_answers():
    m, q = oracle.species_props(3)
    tau = oracle.cooling_time(abs(q), m, 1.0)
    assert tau == pytest.approx(2.579, rel=1e-3)  # README: "about 2.6 s" at 1 T
    assert 2 * math.pi * m / (abs(q) * 1.0) == pytest.approx(35.72e-12, rel=1e-3)  # cyclotron period


def test_cooling_decay_without_bath():
    m, q = oracle.species_props(3)
    vel = np.array([[3e4, -4e4, 1e4]])
    dt = 1e-3
    oracle.cyclotron_cooling(vel, np.array([q]), np.array([m]), np.zeros(1, np.uint8), 1.0, 1.0, 0.0, dt, 1, 1)
    decay = math.exp(-dt / (2 * oracle.cooling_time(abs(q), m, 1.0)))
    np.testing.assert_allclose(vel[0], [3e4 * decay, -4e4 * decay, 1e4], rtol=1e-14)


def test_cooling_noise_is_perpendicular_with_thermal_variance():
    m, q = oracle.species_props(3)
    n = 4000
    vel = np.zeros((n, 3))
    dt = 0.05
    oracle.cyclotron_cooling(vel, np.full(n, q), np.full(n, m), np.zeros(n, np.uint8), 1.0, 1.0, 4.0, dt, 9, 1)
    decay = math.exp(-dt / (2 * oracle.cooling_time(abs(q), m, 1.0)))
    var = oracle.K_BOLTZMANN * 4.0 / m * (1 - decay ** 2)
    assert np.all(vel[:, 2] == 0.0)
    assert vel[:, :2].var() == pytest.approx(var, rel=0.06)


# ---- pair forces --------------------------------------------------------------------------------
def test_two_body_force_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [3e-6, 4e-6, 0.0]])
    q = np.array([E, -E])
    f = oracle.allpairs_forces(pos, q)
    r = 5e-6
    mag = K * E * E / r ** 2
    np.testing.assert_allclose(f[0], mag * np.array([0.6, 0.8, 0.0]), rtol=1e-14)  # attractive
    np.testing.assert_allclose(f[1], -f[0], rtol=1e-15)


def test_allpairs_newton_third_law_and_skips():
    pos, q = cloud(300)
    q[5] = 0.0
    ann = np.zeros(300, np.uint8)
    ann[7] = 1
    f = oracle.allpairs_forces(pos, q, ann)
    assert np.all(f[5] == 0) and np.all(f[7] == 0)
    tot = f.sum(0)
    assert np.linalg.norm(tot) < 1e-12 * np.abs(f).sum()
    # Brute force in numpy for one particle.
    i = 11
    mask = (np.arange(300) != i) & (q != 0) & (ann == 0)
    r = pos[i] - pos[mask]
    want = (K * q[i] * q[mask] / (r ** 2).sum(1) ** 1.5)[:, None] * r
    np.testing.assert_allclose(f[i], want.sum(0), rtol=1e-12)


def test_standard_forces_equal_allpairs_beyond_cloud_and_zero_for_nonpositive_cutoff():
    pos, q = cloud(200, radii=(1e-4,) * 3)
    np.testing.assert_allclose(oracle.standard_forces(pos, q, None, 1.0), oracle.allpairs_forces(pos, q),
                               rtol=1e-12, atol=1e-30)
    assert np.all(oracle.standard_forces(pos, q, None, 0.0) == 0)
    # A cutoff includes exactly the pairs with r^2 < rc^2.
    rc = 3e-5
    f = oracle.standard_forces(pos, q, None, rc)
    i = 3
    r = pos[i] - pos
    d2 = (r ** 2).sum(1)
    m = (d2 < rc * rc) & (np.arange(200) != i)
    want = (K * q[i] * q[m] / d2[m] ** 1.5)[:, None] * r[m]
    np.testing.assert_allclose(f[i], want.sum(0), rtol=1e-12, atol=1e-40)


def test_standard_energy_two_particles_shifted():
    pos = np.array([[0, 0, 0], [2e-6, 0, 0], [0, 0, 5e-5]], dtype=np.float64)
    q = np.array([E, -E, E])
    rc = 1e-5
    want = K * E * (-E) * (1 / 2e-6 - 1 / rc)  # the third particle is beyond the cutoff
    assert oracle.standard_energy(pos, q, None, rc) == pytest.approx(want, rel=1e-14)


def test_allpairs_potential_energy_closed_form():
    pos = np.array([[0, 0, 0], [1e-5, 0, 0], [0, 2e-5, 0]], dtype=np.float64)
    q = np.array([E, -E, E])
    want = K * E * E * (-1 / 1e-5 + 1 / 2e-5 - 1 / math.hypot(1e-5, 2e-5))
    assert oracle.allpairs_potential_energy(pos, q) == pytest.approx(want, rel=1e-14)


# ---- Barnes-Hut --------------------------------------------------------------------------------
def test_bh_theta_zero_is_allpairs():
    pos, q = cloud(1500, seed=3)
    np.testing.assert_allclose(oracle.bh_forces(pos, q, None, 0.0), oracle.allpairs_forces(pos, q), rtol=1e-10,
                               atol=1e-12 * np.abs(oracle.allpairs_forces(pos, q)).max())
    assert oracle.bh_potential_energy(pos, q, None, 0.0) == pytest.approx(oracle.allpairs_potential_energy(pos, q),
                                                                           rel=1e-12)


def test_bh_error_grows_with_theta():
    pos, q = cloud(3000, seed=4)
    ref = oracle.allpairs_forces(pos, q)
    errs = []
    for th in (0.2, 0.5, 0.9):
        f = oracle.bh_forces(pos, q, None, th)
        errs.append(np.linalg.norm(f - ref) / np.linalg.norm(ref))
    assert errs[0] < errs[1] < errs[2] < 0.2
    assert errs[0] < 2e-3


def _partition(pos, idx, c, h, depth, cap, out):
    """Independent recursive restatement of the reference partition (leaf ranges)."""
    if len(idx) <= cap or depth >= 32:
        out.append((depth, list(idx)))
        return
    octs = ((pos[idx, 0] >= c[0]).astype(int) | ((pos[idx, 1] >= c[1]).astype(int) << 1)
            | ((pos[idx, 2] >= c[2]).astype(int) << 2))
    for o in range(8):
        sub = idx[octs == o]
        if len(sub) == 0:
            continue
        ch = 0.5 * h
        cc = (c[0] + (ch if o & 1 else -ch), c[1] + (ch if o & 2 else -ch), c[2] + (ch if o & 4 else -ch))
        _partition(pos, sub, cc, ch, depth + 1, cap, out)


def test_bh_tree_leaf_membership_and_charges():
    pos, q = cloud(2000, seed=5)
    t = oracle.bh_tree(pos, q, 8)
    lo, hi = pos.min(0), pos.max(0)
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo).max() * (1 + 1e-9)
    leaves = []
    _partition(pos, np.arange(len(q)), c, h, 0, 8, leaves)
    got = [list(t["order"][b:e]) for b, e, lf in zip(t["begin"], t["end"], t["leaf"]) if lf]
    assert got == [lv for _, lv in leaves]
    root_qp = q[q > 0].sum()
    assert t["qp"][0] == pytest.approx(root_qp, rel=1e-12)
    assert t["qn"][0] == pytest.approx(-q[q < 0].sum(), rel=1e-12)
    # charge centres lie inside their node cubes
    inside = np.abs(t["cp"] - t["center"]) <= t["half"][:, None] * (1 + 1e-12)
    assert inside[t["qp"] > 0].all()


# ---- FMM ------------------------------------------------------------------------------------------
def test_fmm_derivative_closed_forms():
    r = np.array([0.3, -0.4, 1.2])
    D, multi = oracle.fmm_derivatives(3, r)
    x, y, z = r
    rr = np.linalg.norm(r)
    idx = {tuple(m): k for k, m in enumerate(multi)}
    assert D[idx[(0, 0, 0)]] == pytest.approx(1 / rr, rel=1e-15)
    assert D[idx[(1, 0, 0)]] == pytest.approx(-x / rr ** 3, rel=1e-14)
    assert D[idx[(0, 0, 1)]] == pytest.approx(-z / rr ** 3, rel=1e-14)
    assert D[idx[(2, 0, 0)]] == pytest.approx((3 * x * x - rr * rr) / rr ** 5, rel=1e-13)
    assert D[idx[(1, 1, 0)]] == pytest.approx(3 * x * y / rr ** 5, rel=1e-13)
    assert D[idx[(1, 1, 1)]] == pytest.approx(-15 * x * y * z / rr ** 7, rel=1e-13)
    # Laplace: D_200 + D_020 + D_002 = 0
    assert D[idx[(2, 0, 0)]] + D[idx[(0, 2, 0)]] + D[idx[(0, 0, 2)]] == pytest.approx(0.0, abs=1e-12)


def test_fmm_theta_zero_is_allpairs_and_converges_in_order():
    pos, q = cloud(2500, seed=6)
    ref = oracle.allpairs_forces(pos, q)
    f0 = oracle.fmm_forces(pos, q, None, 4, 0.0)
    assert np.linalg.norm(f0 - ref) / np.linalg.norm(ref) < 1e-12
    errs = [np.linalg.norm(oracle.fmm_forces(pos, q, None, p, 0.5) - ref) / np.linalg.norm(ref) for p in (2, 4, 6)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[1] < 2e-3
    pe = oracle.fmm_potential_energy(pos, q, None, 6, 0.3)
    assert pe == pytest.approx(oracle.allpairs_potential_energy(pos, q), rel=1e-4)


# ---- annihilation --------------------------------------------------------------------------------
def test_annihilation_p1_is_greedy_maximal_matching():
    rng = np.random.default_rng(8)
    n = 400
    pos = rng.uniform(-1, 1, (n, 3)) * 5e-5
    sp = rng.choice([1, 3, 0, 4, 2], n).astype(np.uint8)
    ann = np.zeros(n, np.uint8)
    ann[::17] = 1
    before = ann.copy()
    ra = 1.2e-5
    cnt, pairs = oracle.annihilate(pos, sp, ann, ra, 1.0, 5, 1)
    assert cnt == 2 * len(pairs)
    used = set()
    for i, j in pairs:
        assert oracle.lib().orc_annihilation_uniform(5, 1, int(i), int(j)) < 1.0
        assert i not in used and j not in used
        used |= {int(i), int(j)}
        assert before[i] == 0 and before[j] == 0
        assert np.sum((pos[i] - pos[j]) ** 2) < ra * ra
        assert {int(sp[i]), int(sp[j])} in ({1, 3}, {0, 4})
    # Maximality: no remaining eligible pair among survivors.
    alive = np.where(ann == 0)[0]
    for a in alive:
        for b in alive:
            if a < b and {int(sp[a]), int(sp[b])} in ({1, 3}, {0, 4}):
                assert np.sum((pos[a] - pos[b]) ** 2) >= ra * ra
    # Greedy order: i ascending scans, lowest-index eligible partner.
    assert all(pairs[k][0] < pairs[k + 1][0] for k in range(len(pairs) - 1))


def test_annihilation_probability_half_uses_counter_draws():
    rng = np.random.default_rng(9)
    n = 300
    pos = rng.uniform(-1, 1, (n, 3)) * 3e-5
    sp = np.where(np.arange(n) % 2 == 0, 1, 3).astype(np.uint8)
    a1 = np.zeros(n, np.uint8)
    a2 = np.zeros(n, np.uint8)
    c1, p1 = oracle.annihilate(pos, sp, a1, 1e-5, 0.5, 11, 4)
    c2, p2 = oracle.annihilate(pos, sp, a2, 1e-5, 0.5, 11, 4)
    assert c1 == c2 and np.array_equal(p1, p2)
    for i, j in p1:
        assert oracle.annihilation_uniform(11, 4, int(i), int(j)) < 0.5
