"""Batched ensemble: every trap against an independent single-trap computation."""
import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu


def small_ensemble(ctx, annihilation=False, sizes=(300, 1000, 513), seed=5, radii=None, probability=1.0):
    rng = np.random.default_rng(seed)
    pos_l, vel_l, sp_l, traps, prm = [], [], [], [], []
    b = 0
    for t, n in enumerate(sizes):
        pops = {"positron": n // 2, "electron": n - n // 2} if t % 2 == 0 else {"antiproton": n // 3,
                                                                                "proton": n - n // 3}
        r = radii or tuple(rng.uniform(1e-4, 4e-4, 3))
        pos, vel, sp = am.initial_conditions(pops, {k: 300.0 for k in pops}, r, 10 + t)
        b0, u0 = float(rng.uniform(1, 5)), float(rng.uniform(5, 100))
        f = am.make_trap(b0, u0, 0.01, 0.01)
        dt = am.derive_dt(pops, b0)
        traps.append(am.TrapDesc(b, b + n, f.b0, f.trap_voltage, f.d, dt, 1e-6, 4.0, int(annihilation), 5e-6,
                                 probability))
        prm.append((b, b + n, f, dt))
        pos_l.append(pos); vel_l.append(vel); sp_l.append(sp)
        b += n
    pos, vel, sp = np.concatenate(pos_l), np.concatenate(vel_l), np.concatenate(sp_l)
    s = am.Store(ctx, len(sp))
    s.upload(pos, vel, sp)
    return s, pos, vel, sp, traps, prm


def test_one_step_matches_oracle_per_trap(ctx):
    s, pos, vel, sp, traps, prm = small_ensemble(ctx)
    ens = am.Ensemble(s, traps)
    ens.run(2, seed=77, use_graph=False)
    g = s.download()
    m, q = oracle.mass_charge(sp)
    for t, (b, e, f, dt) in enumerate(prm):
        runs = []
        for coulomb in (True, False):
            p, v = pos[b:e].copy(), vel[b:e].copy()
            st = np.zeros(e - b, np.uint8)
            ann = np.zeros(e - b, np.uint8)
            for step in (1, 2):
                fc = oracle.allpairs_forces(p, q[b:e]) if coulomb else np.zeros_like(p)
                oracle.lorentz_push(p, v, fc, q[b:e], m[b:e], ann, st, f.b0, f.trap_voltage, f.d, dt)
                oracle.cyclotron_cooling(v, q[b:e], m[b:e], ann, f.b0, 1e-6, 4.0, dt, am.trap_seed(77, t), step)
            runs.append((p, v))
        (p, v), (p0, v0) = runs
        # FP32 pair forces: the error is a small fraction of the Coulomb part of the motion (the
        # cooling noise and the trap motion are reproduced to rounding).
        dx_coul, dv_coul = np.abs(p - p0).max(), np.abs(v - v0).max()
        assert np.abs(g["pos"][b:e] - p).max() <= 1e-5 * dx_coul + 1e-18
        assert np.abs(g["vel"][b:e] - v).max() <= 1e-5 * dv_coul + 1e-9 * np.abs(v).max()
    ens.close()
    s.close()


def test_graph_replay_equals_eager_and_chunking_is_stable(ctx):
    outs = []
    for use_graph, chunk, ann in ((False, 4096, True), (True, 4096, True), (True, 4096, False), (True, 256, False)):
        s, pos, vel, sp, traps, prm = small_ensemble(ctx, annihilation=ann, radii=(3e-5, 3e-5, 3e-5))
        ens = am.Ensemble(s, traps)
        ens.set_chunk(chunk)
        ens.run(3, seed=9, use_graph=use_graph)
        outs.append((s.download(), ens.stats()))
        ens.close()
        s.close()
    a, b, c, d = outs
    np.testing.assert_array_equal(a[0]["pos"], b[0]["pos"])
    np.testing.assert_array_equal(a[0]["annihilated"], b[0]["annihilated"])
    assert a[1]["annihilated"] == b[1]["annihilated"] > 0
    # a different source chunking only reorders FP32 partial sums
    np.testing.assert_allclose(d[0]["pos"], c[0]["pos"], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("p", [1.0, 0.5])
def test_annihilation_matches_oracle_per_trap(ctx, p):
    s, pos, vel, sp, traps, prm = small_ensemble(ctx, annihilation=True, radii=(3e-5, 3e-5, 3e-5), probability=p)
    ens = am.Ensemble(s, traps)
    ens.run(1, seed=123, use_graph=True)
    g = s.download()
    total = 0
    for t, (b, e, f, dt) in enumerate(prm):
        ann = np.zeros(e - b, np.uint8)
        cnt, _ = oracle.annihilate(g["pos"][b:e].copy(), sp[b:e].copy(), ann, 5e-6, p, am.trap_seed(123, t), 1)
        np.testing.assert_array_equal(g["annihilated"][b:e], ann)
        total += cnt
    assert total > 0 and ens.stats()["annihilated"] == total
    ens.close()
    s.close()


def test_autotuner_picks_a_candidate(ctx):
    s, pos, vel, sp, traps, prm = small_ensemble(ctx, sizes=(4000, 1500, 700, 2500))
    ens = am.Ensemble(s, traps)
    chosen, times = ens.autotune([256, 1024, 4096])
    assert chosen in (256, 1024, 4096) and np.all(times > 0)
    assert ens.stats()["chunk"] == chosen
    ens.close()
    s.close()


def test_invalid_trap_ranges(ctx):
    s, pos, vel, sp, traps, prm = small_ensemble(ctx)
    bad = list(traps)
    bad[1] = am.TrapDesc(100, 200, 1, 1, 0.01, 1e-12, 0, 0, 0, 0, 0)  # overlaps trap 0
    with pytest.raises(am.AmError):
        am.Ensemble(s, bad)
    s.close()
