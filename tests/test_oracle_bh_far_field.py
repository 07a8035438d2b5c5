"""Pins the oracle's Barnes-Hut far field at theta > 0 (barnes-hut.cc:183-221, 223-269).

A hand-built cloud makes the far-field decisions of the walk known in advance:

* target T (+e) at the origin;
* a cluster of 4 positive and 4 negative unit charges inside [0.51, 0.60]^3 (units L = 1e-4 m);
* a neutral particle at (1, 1, 1), which sets the bounding box to [0, 1]^3.

With the reference's cube (centre = bbox mid, half = ext/2 (1 + 1e-9)), the cluster and the
neutral particle share the root's octant 7: a cube centred at (0.75, 0.75, 0.75) with half
width 0.25 holding 9 particles.  It is an internal node at leaf capacity 8, and its children
(octant 0 with the 8 charges, octant 7 with the neutral one) are leaves.  Seen from T:

* the geometric centre is at distance 0.75 sqrt(3), so the MAC (2h)^2 < theta^2 d^2 accepts the
  node for theta > 0.385 and opens it below;
* both class charge centres lie near (0.55, 0.55, 0.55), so a MAC measured from either of them
  (or from the mean of the charges) opens it for every theta below 0.52;
* a width of h instead of 2h accepts it for theta > 0.19.

So at theta = 0.45 the reference walk adds the two class monopoles (Q+ at c+, -Q- at c-) and at
theta = 0.35 it sums the 8 charges directly.  Both expectations are computed here from the
Coulomb pair term and the class sums alone.  They differ by 4 % in force and 9 % in potential,
far above rounding, so a MAC from a charge centre, a width of h, or a dropped class monopole
gives a result that misses them; ``test_planted_mistakes_are_caught`` compiles the oracle with
each of these mistakes planted and checks that the scenario rejects it.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

E = oracle.E_CHARGE
K = oracle.K_COULOMB
L = 1e-4
ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "cpu_ref.cpp")


def scenario():
    cluster = np.array([[0.52, 0.55, 0.53], [0.58, 0.51, 0.56], [0.54, 0.59, 0.60], [0.60, 0.53, 0.51],
                        [0.51, 0.57, 0.58], [0.57, 0.60, 0.52], [0.55, 0.52, 0.59], [0.59, 0.58, 0.55],
                        [1.0, 1.0, 1.0]])
    qc = np.array([1, -1, 1, -1, 1, -1, 1, -1, 0.0])
    pos = np.vstack([np.zeros(3), cluster]) * L
    q = np.concatenate([[1.0], qc]) * E
    return pos, q


def _pair_force(t, qt, s, qs):
    r = t - s
    return K * qt * qs * r / (r @ r) ** 1.5


def expectations():
    """Force and potential on T: (exact direct sum, dual-monopole of the cluster node)."""
    pos, q = scenario()
    t = pos[0]
    src = [k for k in range(1, len(q)) if q[k] != 0.0]
    f_exact = sum(_pair_force(t, q[0], pos[k], q[k]) for k in src)
    phi_exact = sum(q[k] / np.linalg.norm(t - pos[k]) for k in src)
    pp = [k for k in src if q[k] > 0]
    nn = [k for k in src if q[k] < 0]
    qp = q[pp].sum()
    qn = -q[nn].sum()
    cp = (q[pp, None] * pos[pp]).sum(0) / qp
    cn = (-q[nn, None] * pos[nn]).sum(0) / qn
    f_mono = _pair_force(t, q[0], cp, qp) + _pair_force(t, q[0], cn, -qn)
    phi_mono = qp / np.linalg.norm(t - cp) - qn / np.linalg.norm(t - cn)
    return dict(f_exact=f_exact, phi_exact=phi_exact, f_mono=f_mono, phi_mono=phi_mono, cp=cp, cn=cn,
                f_pos_only=_pair_force(t, q[0], cp, qp))


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


class _OracleLib:
    """The two oracle functions the scenario needs, from a given liboracle build."""

    def __init__(self, path):
        self.L = ctypes.CDLL(path)
        vp, d, i64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64
        self.L.orc_bh_forces.argtypes = [i64, vp, vp, vp, d, i64, vp, vp]
        self.L.orc_bh_forces.restype = None
        self.L.orc_bh_potentials.argtypes = [i64, vp, vp, vp, d, i64, vp]
        self.L.orc_bh_potentials.restype = None

    def forces(self, pos, q, theta):
        f = np.zeros_like(pos)
        self.L.orc_bh_forces(len(q), pos.ctypes.data, q.ctypes.data, None, theta, 8, f.ctypes.data, None)
        return f

    def potentials(self, pos, q, theta):
        phi = np.zeros(len(q))
        self.L.orc_bh_potentials(len(q), pos.ctypes.data, q.ctypes.data, None, theta, 8, phi.ctypes.data)
        return phi


def scenario_failures(forces, potentials):
    """Names of the scenario's expectations that a Barnes-Hut implementation misses."""
    pos, q = scenario()
    ex = expectations()
    bad = []
    if _rel(forces(pos, q, 0.45)[0], ex["f_mono"]) > 1e-12:
        bad.append("force theta=0.45 (monopoles)")
    if _rel(forces(pos, q, 0.35)[0], ex["f_exact"]) > 1e-12:
        bad.append("force theta=0.35 (direct)")
    if abs(potentials(pos, q, 0.45)[0] / ex["phi_mono"] - 1) > 1e-12:
        bad.append("potential theta=0.45 (monopoles)")
    if abs(potentials(pos, q, 0.35)[0] / ex["phi_exact"] - 1) > 1e-12:
        bad.append("potential theta=0.35 (direct)")
    return bad


def test_scenario_premises():
    """The geometry does what the module docstring says (checked from the cube rule, not the oracle)."""
    pos, q = scenario()
    lo, hi = pos.min(0), pos.max(0)
    np.testing.assert_allclose(lo, 0.0)
    np.testing.assert_allclose(hi, L)
    centre = 0.5 * (lo + hi) + 0.25 * L  # the root's octant 7
    half = 0.25 * L * (1 + 1e-9)
    assert np.all(pos[1:] >= 0.5 * L)
    assert np.all(pos[1:9] < centre)  # the 8 charges share the child octant 0: a leaf
    ex = expectations()
    t = pos[0]

    def accepts(c, width, theta):
        return width ** 2 < theta ** 2 * float(((t - c) ** 2).sum())

    assert accepts(centre, 2 * half, 0.45) and not accepts(centre, 2 * half, 0.35)
    for c in (ex["cp"], ex["cn"], pos[1:9].mean(0)):
        assert not accepts(c, 2 * half, 0.45)
    assert accepts(centre, half, 0.35)
    # the two outcomes are far apart, and each class monopole matters
    assert _rel(ex["f_mono"], ex["f_exact"]) > 0.03
    assert abs(ex["phi_mono"] / ex["phi_exact"] - 1) > 0.05
    assert _rel(ex["f_pos_only"], ex["f_mono"]) > 1.0


def test_oracle_far_field_matches_hand_sums():
    lib = _OracleLib(os.path.join(os.path.dirname(ORACLE_SRC), "liboracle.so"))
    oracle.lib()  # builds liboracle.so if needed
    assert scenario_failures(lib.forces, lib.potentials) == []


def test_oracle_python_binding_on_scenario():
    pos, q = scenario()
    ex = expectations()
    assert _rel(oracle.bh_forces(pos, q, None, 0.45)[0], ex["f_mono"]) < 1e-12
    assert oracle.bh_potentials(pos, q, None, 0.35)[0] == pytest.approx(ex["phi_exact"], rel=1e-12)
    # the energy walk uses the same potential: U = k/2 sum_i q_i phi_i
    phi = oracle.bh_potentials(pos, q, None, 0.45)
    assert oracle.bh_potential_energy(pos, q, None, 0.45) == pytest.approx(0.5 * K * float(q @ phi), rel=1e-12)


def test_oracle_evaluate_at_points():
    """BH evaluation at given points: at the particles themselves (self = own index) it is the pinned
    force and potential walk; at theta = 0 it is the direct sum at arbitrary points."""
    rng = np.random.default_rng(9)
    n = 800
    pos = rng.uniform(-1, 1, (n, 3)) * 1e-4
    q = np.where(rng.random(n) < 0.5, E, -E)
    ann = (rng.random(n) < 0.05).astype(np.uint8)
    f, phi = oracle.bh_evaluate(pos, q, ann, 0.6, pos, q, np.arange(n))
    live = ann == 0
    np.testing.assert_array_equal(f[live], oracle.bh_forces(pos, q, ann, 0.6)[live])
    np.testing.assert_array_equal(phi[live], oracle.bh_potentials(pos, q, ann, 0.6)[live])
    pts = rng.uniform(-2, 2, (50, 3)) * 1e-4
    qt = np.full(50, -E)
    f0, phi0 = oracle.bh_evaluate(pos, q, ann, 0.0, pts, qt, None)
    r = pts[:, None, :] - pos[None, live, :]
    d = np.linalg.norm(r, axis=2)
    want_f = (K * qt[:, None, None] * q[None, live, None] * r / d[..., None] ** 3).sum(1)
    want_phi = (q[None, live] / d).sum(1)
    np.testing.assert_allclose(f0, want_f, rtol=1e-12, atol=1e-12 * np.abs(want_f).max())
    np.testing.assert_allclose(phi0, want_phi, rtol=1e-12, atol=1e-12 * np.abs(want_phi).max())
    # self leaves exactly that source out, wherever the point is
    f1, phi1 = oracle.bh_evaluate(pos, q, ann, 0.0, pts[:1], qt[:1], np.array([3]))
    np.testing.assert_allclose(phi1[0], want_phi[0] - q[3] / d[0, np.where(live)[0] == 3][0], rtol=1e-12)


# Mistakes planted in a copy of the oracle source: (name, [(text, replacement), ...]).  Each text
# must occur in the source; every occurrence is replaced (the force and the potential walks).
MUTANTS = [
    ("mac_from_positive_charge_centre", [("const V3 d = sub(x, nd.center);", "const V3 d = sub(x, nd.cp);")]),
    ("mac_from_negative_charge_centre", [("const V3 d = sub(x, nd.center);", "const V3 d = sub(x, nd.cn);")]),
    ("width_is_half", [("const double w = 2.0 * nd.half;", "const double w = nd.half;")]),
    ("negative_monopole_dropped", [("if (nd.qn > 0.0) { PointCharge(x, qt, nd.cn, -nd.qn, f); ++*count; }", ""),
                                   ("if (r2 >= kMinSq) *phi -= nd.qn / std::sqrt(r2);", "(void)r2;")]),
    ("positive_monopole_dropped", [("if (nd.qp > 0.0) { PointCharge(x, qt, nd.cp, nd.qp, f); ++*count; }", ""),
                                   ("if (r2 >= kMinSq) *phi += nd.qp / std::sqrt(r2);", "(void)r2;")]),
]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the planted-mistake oracles")
@pytest.mark.parametrize("name,edits", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistakes_are_caught(tmp_path, name, edits):
    with open(ORACLE_SRC) as fh:
        src = fh.read()
    for text, repl in edits:
        assert text in src, f"mutation site for {name} not found in the oracle: {text}"
        src = src.replace(text, repl)
    cpp = tmp_path / f"{name}.cpp"
    so = tmp_path / f"lib_{name}.so"
    cpp.write_text(src)
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o", str(so),
                    str(cpp)], check=True)
    lib = _OracleLib(str(so))
    bad = scenario_failures(lib.forces, lib.potentials)
    # every planted mistake changes the force and the potential walk alike
    assert any(b.startswith("force") for b in bad), f"{name}: force expectations still met"
    assert any(b.startswith("potential") for b in bad), f"{name}: potential expectations still met"
