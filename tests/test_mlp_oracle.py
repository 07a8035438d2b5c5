"""Pins of the learned-correction oracle (oracle/mlp_ref.py): bf16 rounding against known values,
the hand-written backward against finite differences, and the residual/feature restatement
against the multipole identities."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def test_bf16_round_known_values():
    cases = [
        (1.0, 1.0),
        (1.0 + 2.0 ** -8, 1.0),                       # tie, even neighbour below
        (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6),       # tie, even neighbour above
        (1.0 + 2.0 ** -8 + 2.0 ** -20, 1.0 + 2.0 ** -7),  # just above the tie
        (-(1.0 + 2.0 ** -8 + 2.0 ** -20), -(1.0 + 2.0 ** -7)),
        (3.0e38, 226.0 * 2.0 ** 120),                 # 3e38 / 2^120 = 225.7
        (3.4028234663852886e38, np.inf),              # past the largest bf16 + half an ulp
        (3.3895313892515355e38, 3.3895313892515355e38),
        (2.0 ** -133, 2.0 ** -133),                   # subnormals keep their top 7 bits
        (0.0, 0.0),
    ]
    for x, want in cases:
        got = M.bf16_round(np.float32(x))
        assert got == want, (x, got, want)
    assert np.isnan(M.bf16_round(np.float32(np.nan)))
    # bf16 keeps 8 significant bits: results are representable with 8 bits and the relative
    # rounding error is at most 2^-8 on normal values
    x = np.random.default_rng(0).normal(size=10000).astype(np.float32)
    r = M.bf16_round(x)
    mant, _ = np.frexp(r)
    assert np.all(mant * 256 == np.round(mant * 256))
    assert np.all(np.abs(r - x) <= 2.0 ** -8 * np.abs(x))


def _net(seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros(M.N_PARAMS)
    p = M.unpack(w)
    p["W1"] = rng.normal(0, 0.5, p["W1"].shape)
    p["b1"] = rng.normal(0, 0.1, p["b1"].shape)
    p["W2"] = rng.normal(0, 0.2, p["W2"].shape)
    p["b2"] = rng.normal(0, 0.1, p["b2"].shape)
    p["W3"] = rng.normal(0, 0.2, p["W3"].shape)
    p["b3"] = rng.normal(0, 0.1, p["b3"].shape)
    return M.pack(p)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = _net()
    X = rng.normal(size=(7, M.N_IN))
    Y = rng.normal(size=(7, M.N_OUT))
    loss, g = M.loss_and_grad(w, X, Y, rounding=False)
    idx = rng.choice(M.N_PARAMS, 200, replace=False)
    # include every bias block and the corners of each weight block
    idx = np.unique(np.concatenate([idx, [0, 16 * 64 - 1, 16 * 64, 16 * 64 + 63, 16 * 64 + 64, M.N_PARAMS - 1]]))
    eps = 1e-6
    fd = np.empty(len(idx))
    for k, p in enumerate(idx):
        wp, wm = w.copy(), w.copy()
        wp[p] += eps
        wm[p] -= eps
        fd[k] = (M.loss_and_grad(wp, X, Y, False)[0] - M.loss_and_grad(wm, X, Y, False)[0]) / (2 * eps)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())
    assert loss == pytest.approx(float(((M.forward(w, X, False)[0] - Y) ** 2).mean()))


def test_rounded_forward_differs_from_unrounded_by_bf16_precision():
    rng = np.random.default_rng(2)
    w = _net(3)
    X = rng.normal(size=(64, M.N_IN))
    y0 = M.forward(w, X, rounding=False)[0]
    y1 = M.forward(w, X, rounding=True)[0]
    rel = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 1e-4 < rel < 3e-2


def _cluster(rng, n, half, centre, sign_mix=True):
    pos = centre + rng.uniform(-half, half, (n, 3))
    q = rng.choice([-1.0, 1.0], n) if sign_mix else np.ones(n)
    return pos, q


def test_residual_vanishes_for_coincident_charges():
    pos = np.tile([[1.0, 2.0, 3.0]], (5, 1))
    q = np.array([1.0, 1.0, -1.0, 1.0, -1.0])
    for sign in (0, 1):
        r = M.node_residual(np.array([50.0, -20.0, 7.0]), pos, q, sign)
        assert np.all(np.abs(r) < 1e-18)


def test_quadrupole_is_leading_term_of_residual():
    rng = np.random.default_rng(4)
    pos, q = _cluster(rng, 40, 1.0, np.zeros(3))
    direction = np.array([0.3, -0.5, 0.81])
    direction /= np.linalg.norm(direction)
    for sign in (0, 1):
        errs = []
        for R in (10.0, 20.0, 40.0, 80.0):
            t = R * direction
            res = M.node_residual(t, pos, q, sign)
            quad = M.quadrupole_residual(t, pos, q, sign)
            errs.append(np.linalg.norm(res - quad) / np.linalg.norm(quad))
        # relative error of the quadrupole term is O(h/R): halves as R doubles
        slopes = np.diff(np.log(errs)) / np.log(2.0)
        assert errs[0] < 0.5
        assert np.all(np.abs(slopes + 1.0) < 0.3), slopes


def test_features_are_dimensionless_and_bounded():
    rng = np.random.default_rng(5)
    centre = np.array([3.0, -1.0, 2.0])
    half = 0.5
    pos, q = _cluster(rng, 30, half, centre)
    t = centre + np.array([4.0, 1.0, -2.0])
    for sign in (0, 1):
        f, Q, R = M.node_features(t, pos, q, sign, centre, half)
        assert np.linalg.norm(f[0:3]) == pytest.approx(1.0)
        assert f[3] == pytest.approx(half / R)
        assert np.all(np.abs(f[4:7]) <= 1.0)
        assert np.all(f[7:10] >= 0) and np.all(f[7:10] <= 1.0)  # per-axis variance <= h^2
        assert np.all(f[13:] == 0)
        # scaling all lengths leaves the features unchanged
        g, _, _ = M.node_features(2.5 * t, 2.5 * pos, q, sign, 2.5 * centre, 2.5 * half)
        np.testing.assert_allclose(g, f, rtol=1e-12, atol=1e-14)


def test_quadrupole_from_features_matches_particle_sum():
    rng = np.random.default_rng(6)
    centre = np.array([0.5, 0.2, -0.3])
    half = 1.0
    pos, q = _cluster(rng, 25, half, centre)
    for sign in (0, 1):
        sigma = 1.0 if sign == 0 else -1.0
        for t in (np.array([6.0, 2.0, -1.0]), np.array([-3.0, 4.0, 5.0])):
            f, Q, R = M.node_features(t, pos, q, sign, centre, half)
            y = M.quadrupole_from_features(f[None, :])[0] * sigma * Q * half ** 2 / R ** 4
            np.testing.assert_allclose(y, M.quadrupole_residual(t, pos, q, sign), rtol=1e-12)


def test_product_quadrupole_baseline_matches_oracle():
    """am_mlp_quadrupole_baseline is host code of the library: FP32 against the oracle's FP64."""
    from antimatter_amd import mlp
    rng = np.random.default_rng(7)
    X = np.zeros((500, 16), np.float32)
    u = rng.normal(size=(500, 3))
    X[:, 0:3] = u / np.linalg.norm(u, axis=1, keepdims=True)
    X[:, 3] = rng.uniform(0.05, 0.5, 500)
    X[:, 4:7] = rng.uniform(-1, 1, (500, 3))
    X[:, 7:10] = rng.uniform(0, 1, (500, 3))
    X[:, 10:13] = rng.uniform(-0.3, 0.3, (500, 3))
    y = mlp.quadrupole_baseline(X)
    ref = M.quadrupole_from_features(X)
    np.testing.assert_allclose(y, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
