"""The seeded device cloud generator (am_store_generate_cloud) by its statistics: it draws the
reference's distributions (penning_trap.cc:51-161) with counter-based draws, so only the laws are
checked (each bound is 5 standard errors of the estimate), plus the exact properties the reference's
construction fixes: every position inside the ellipsoid, species in enum-ordered blocks, each
species' mean velocity zero, and a cloud that depends only on the seed."""
import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu

KB = oracle.K_BOLTZMANN


def generate(ctx, pops, temps, radii, seed, cap=None):
    n = sum(pops.values())
    s = am.Store(ctx, cap or n)
    s.generate_cloud(pops, temps, radii, seed)
    return s


def test_config3_size_cloud_statistics(ctx):
    n = 1 << 20
    pops = {"positron": n // 2, "electron": n // 2}
    radii = np.array([2e-3, 2e-3, 8e-3])
    print(f"icgen: n={n}", flush=True)
    s = generate(ctx, pops, {"positron": 4.0, "electron": 4.0}, radii, 42)
    g = s.download()
    u = g["pos"] / radii
    r2 = (u ** 2).sum(1)
    assert np.all(r2 <= 1.0 + 1e-12)
    se = 5 * np.sqrt(8 / 175 / n)  # Var(x^2) = 8/175 for uniform in the unit ball
    assert np.all(np.abs((u ** 2).mean(0) - 0.2) < se)
    assert np.all(np.abs(u.mean(0)) < 5 * np.sqrt(0.2 / n))
    p = 1 / 8  # inside the half-size ellipsoid
    assert abs((r2 < 0.25).mean() - p) < 5 * np.sqrt(p * (1 - p) / n)
    sp = s.species()
    np.testing.assert_array_equal(sp, np.repeat(np.array([1, 3], np.uint8), n // 2))
    m, q = oracle.mass_charge(sp)
    for species, (b, e) in ((1, (0, n // 2)), (3, (n // 2, n))):
        v = g["vel"][b:e]
        sigma2 = KB * 4.0 / m[b]
        assert np.all(np.abs(v.mean(0)) < 1e-12 * np.sqrt(sigma2))
        var = (v ** 2).mean(0) / sigma2
        assert np.all(np.abs(var - 1) < 5 * np.sqrt(2 / (e - b)))
        assert abs((v ** 2).mean() / sigma2 - 1) < 5 * np.sqrt(2 / (3 * (e - b)))
    e = s.compute_total_energy(am.make_trap(1.0, 10.0, 0.01, 0.01), "standard", cutoff_radius=1e-9)
    assert e.kinetic == pytest.approx(1.5 * n * KB * 4.0, rel=5 * np.sqrt(2 / (3 * n)))
    assert e.num_active_particles == n
    s.close()


def test_seed_determinism_and_mixed_species(ctx):
    pops = {"antiproton": 3000, "positron": 5000, "electron": 1, "proton": 700}
    temps = {"antiproton": 10.0, "positron": 100.0, "electron": 50.0, "proton": 0.0}
    radii = (1e-3, 1e-3, 4e-3)
    a = generate(ctx, pops, temps, radii, 7, cap=20000)
    b = generate(ctx, pops, temps, radii, 7)
    c = generate(ctx, pops, temps, radii, 8)
    ga, gb, gc = a.download(), b.download(), c.download()
    np.testing.assert_array_equal(ga["pos"], gb["pos"])
    np.testing.assert_array_equal(ga["vel"], gb["vel"])
    assert not np.array_equal(ga["pos"], gc["pos"])
    sp = a.species()
    np.testing.assert_array_equal(sp, np.repeat(np.array([0, 1, 3, 4], np.uint8), [3000, 5000, 1, 700]))
    v = ga["vel"]
    assert np.all(v[8001:] == 0.0)  # the cold species is at rest
    assert np.all(np.isfinite(v)) and np.all(v[8000] != 0.0)  # a lone particle keeps its draw
    # the generated store runs the solvers: FP64 all-pairs equals the oracle on it
    m, q = oracle.mass_charge(sp)
    a.compute_allpairs_coulomb_forces(am.FP64)
    np.testing.assert_array_equal(a.forces(), oracle.allpairs_forces(ga["pos"], q))
    with pytest.raises(am.AmError):
        generate(ctx, {"positron": 10}, {"positron": 1.0}, (1e-3, 0.0, 1e-3), 1)
    for st in (a, b, c):
        st.close()


def test_ensemble_size_cloud(ctx):
    n = 4_700_000
    pops = {"positron": 1_500_000, "electron": 1_400_000, "antiproton": 1_000_000, "proton": 800_000}
    print(f"icgen: n={n}", flush=True)
    s = generate(ctx, pops, {k: 300.0 for k in pops}, (2e-3, 2e-3, 2e-3), 3)
    assert s.count() == (n, n)
    g = s.download()
    assert np.all(np.isfinite(g["pos"])) and np.all(np.isfinite(g["vel"]))
    s.close()
