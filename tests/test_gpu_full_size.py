"""Parity at the workloads' full sizes and mixes.

* The Barnes-Hut / FMM workload: N = 1,048,576 positrons + electrons (50/50), radii
  (2e-3, 2e-3, 8e-3) m, 4 K.  The octree (order, node ranges, leaves, class charges) is identical
  to the oracle's and needs no depth beyond the 21 levels of the 63-bit keys; BH theta = 0.7 and
  FMM order 4, theta = 0.5 forces on 512 seeded targets equal the oracle's (FP32 paths within
  1e-5 global relative error; FP64 FMM within 1e-9).
* The FMM tile scan at N = 65,536 and theta = 0: every level's frontier holds millions of cell
  pairs (more than one 1024-tile chunk of the scan), and the result is the exact all-pairs sum.
* The batched workload's trap generator with its four-species Dirichlet mix (e+e- and pbar-p
  pairs and non-pairs in one trap's contact lists): per trap, one step of forces, push with
  cooling and annihilation against the oracle.
"""
import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def tree_cloud():
    n = 1 << 20
    pops = {"positron": n // 2, "electron": n - n // 2}
    pos, vel, sp = am.initial_conditions(pops, {"positron": 4.0, "electron": 4.0}, (2e-3, 2e-3, 8e-3), 42)
    m, q = oracle.mass_charge(sp)
    return pos, vel, sp, q


def test_tree_at_one_million_is_the_reference_tree(ctx, tree_cloud):
    pos, vel, sp, q = tree_cloud
    print(f"full-size tree: n={len(sp)}", flush=True)
    s = am.Store(ctx, len(sp))
    s.upload(pos, vel, sp)
    s.compute_barnes_hut_coulomb_forces(0.7, am.FP32)
    t = s.bh_tree()
    ref = oracle.bh_tree(pos, q, 8)
    assert t["deep_overflow"] == 0
    assert t["max_depth"] == int(ref["depth"].max()) <= 21
    np.testing.assert_array_equal(t["order"], ref["order"])
    np.testing.assert_array_equal(t["begin"], ref["begin"])
    np.testing.assert_array_equal(t["end"], ref["end"])
    np.testing.assert_array_equal(t["leaf"], ref["leaf"])
    np.testing.assert_array_equal(t["qp"], ref["qp"])
    np.testing.assert_array_equal(t["qn"], ref["qn"])
    s.close()


def test_bh_and_fmm_forces_at_one_million(ctx, tree_cloud):
    pos, vel, sp, q = tree_cloud
    n = len(sp)
    targets = np.sort(np.random.default_rng(8).choice(n, 512, replace=False))
    s = am.Store(ctx, n)
    s.upload(pos, vel, sp)
    print(f"full-size BH theta=0.7: n={n}", flush=True)
    s.compute_barnes_hut_coulomb_forces(0.7, am.FP32)
    f = s.forces()[targets]
    ref = oracle.bh_forces(pos, q, None, 0.7)[targets]
    assert rel_l2(f, ref) < 1e-5
    print(f"full-size FMM P=4 theta=0.5: n={n}", flush=True)
    ref = oracle.fmm_forces(pos, q, None, 4, 0.5)[targets]
    s.compute_fast_multipole_coulomb_forces(4, 0.5, am.FP64)
    assert rel_l2(s.forces()[targets], ref) < 1e-9
    s.compute_fast_multipole_coulomb_forces(4, 0.5, am.FP32)
    assert rel_l2(s.forces()[targets], ref) < 1e-5
    # the method error itself, against the exact sum on the same targets
    direct = oracle.allpairs_forces_subset(pos, q, None, targets)
    err = rel_l2(s.forces()[targets], direct)
    print(f"FMM P=4 theta=0.5 error vs direct on 512 targets: {err:.3e}")
    assert err < 1e-2
    s.close()


def test_fmm_theta_zero_multi_chunk_scan(ctx):
    rng = np.random.default_rng(17)
    n = 65536
    pos = rng.uniform(-1, 1, (n, 3)) * np.array([1e-3, 1e-3, 4e-3])
    sp = rng.choice([1, 3], n).astype(np.uint8)
    print(f"FMM theta=0: n={n}", flush=True)
    s = am.Store(ctx, n)
    s.upload(pos, np.zeros_like(pos), sp)
    s.compute_fast_multipole_coulomb_forces(4, 0.0, am.FP64)
    f = s.forces()
    st = s.fmm_stats()
    s.compute_allpairs_coulomb_forces(am.FP64)
    ap = s.forces()
    assert st["m2l_pairs"] == 0
    assert st["p2p_cell_pairs"] > 1024 * 1024  # the deepest frontier spans several scan chunks
    assert rel_l2(f, ap) < 1e-12
    s.close()


def test_config5_species_mix_one_step_per_trap(ctx):
    # The workload's generator (4-species Dirichlet mix, log-uniform N and T, B0, U0, cooling at 4 K),
    # with smaller clouds so that contacts occur at r_a = 5e-6 m.
    pos, vel, sp, traps, params = am.make_config5(n_traps=6, seed=7, n_min=1024, n_max=4096,
                                                  radii_range=(4e-5, 8e-5))
    mixes = [len(set(sp[t.begin:t.end].tolist())) for t in traps]
    assert max(mixes) == 4
    s = am.Store(ctx, len(sp))
    s.upload(pos, vel, sp)
    ens = am.Ensemble(s, traps)
    ens.run(1, seed=31, use_graph=True)
    g = s.download()
    m, q = oracle.mass_charge(sp)
    total = 0
    pairs_seen = set()
    for t, (tr, prm) in enumerate(zip(traps, params)):
        b, e = tr.begin, tr.end
        print(f"config5 mix: trap {t} n={e - b} species={sorted(set(sp[b:e].tolist()))}", flush=True)
        runs = []
        for coulomb in (True, False):
            p, v = pos[b:e].copy(), vel[b:e].copy()
            st = np.zeros(e - b, np.uint8)
            ann = np.zeros(e - b, np.uint8)
            fc = oracle.allpairs_forces(p, q[b:e]) if coulomb else np.zeros_like(p)
            oracle.lorentz_push(p, v, fc, q[b:e], m[b:e], ann, st, tr.b0, tr.trap_voltage, tr.d, tr.dt)
            oracle.cyclotron_cooling(v, q[b:e], m[b:e], ann, tr.b0, tr.tau_scale, tr.bath_temp, tr.dt,
                                     am.trap_seed(31, t), 1)
            runs.append((p, v))
        (p, v), (p0, v0) = runs
        dx_coul = np.abs(p - p0).max()
        assert np.abs(g["pos"][b:e] - p).max() <= 1e-5 * dx_coul + 1e-18
        # annihilation on the post-push positions of the device
        ann = np.zeros(e - b, np.uint8)
        cnt, _ = oracle.annihilate(g["pos"][b:e].copy(), sp[b:e].copy(), ann, tr.contact_radius, tr.probability,
                                   am.trap_seed(31, t), 1)
        np.testing.assert_array_equal(g["annihilated"][b:e], ann)
        total += cnt
        gone = sp[b:e][ann == 1]
        pairs_seen |= set(gone.tolist())
    assert total > 0 and ens.stats()["annihilated"] == total
    assert {1, 3} <= pairs_seen and {0, 4} <= pairs_seen  # both pair types annihilated
    ens.close()
    s.close()
