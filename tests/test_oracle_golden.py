"""The CPU oracle against the reference's only published results: the notebook rows of the
shipped params.json run (simulation_analysis.ipynb:161-172,367,438-442,534)."""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


@pytest.fixture(scope="module")
def golden():
    return json.load(open(os.path.join(GOLDEN, "config1_notebook.json")))


@pytest.fixture(scope="module")
def params():
    return json.load(open(os.path.join(GOLDEN, "params_config1.json")))


@pytest.fixture(scope="module")
def run(params):
    return oracle.simulate(params)


def sig6(x):
    return float("%.5e" % x)


def test_rows_0_to_4_match_notebook(run, golden):
    rows, _ = run
    # The CSV holds 6 significant digits; the cooling noise (bath 4 K) is drawn from a different
    # stream than BioDynaMo's and moves the 6th digit of the drift columns at most by one.
    for g in golden["rows"]:
        r = rows[g["step"]]
        assert int(r[0]) == g["step"]
        for k, col in (("kinetic", 1), ("trap_potential", 2), ("energy_potential", 3), ("total", 4)):
            assert sig6(r[col]) == pytest.approx(g[k], rel=1.01e-5), (g["step"], k, r[col], g[k])
        assert int(r[5]) == g["num_active"]
        assert r[6] == pytest.approx(g["fractional_drift"], abs=1.0e-6)
        assert r[7] == pytest.approx(g["absolute_drift"], rel=2.0e-5)


def test_drift_windows_final_and_active(run, golden):
    rows, fin = run
    assert len(rows) == 101  # log interval 1 for 100 steps (antimatter-kernel.cc:108)
    drift = rows[:, 6]
    windows = [drift[k + 20] - drift[k] for k in (0, 20, 40, 60, 80)]
    np.testing.assert_allclose(windows, golden["drift_change_per_20_steps"], atol=1.5e-6)
    assert drift[-1] == pytest.approx(golden["final_fractional_drift"], abs=1.0e-6)
    assert np.abs(drift).max() == pytest.approx(golden["max_abs_fractional_drift"], abs=1.0e-6)
    assert int(rows[0, 5]) == golden["active_start"]
    assert int(rows[-1, 5]) == golden["active_end"]
    assert int((fin["annihilated"] == 0).sum()) == golden["active_end"]


def test_cooling_effect_is_below_print_precision(params):
    """The notebook run has cooling attached (scale 1, bath 4 K); its effect is below 1e-5."""
    on, _ = oracle.simulate(params)
    p = dict(params)
    p["cooling_scale"] = 0
    off, _ = oracle.simulate(p)
    np.testing.assert_allclose(on[:, 1:5], off[:, 1:5], rtol=1e-5)


def test_initial_conditions_order_and_statistics(params):
    pops = params["particle_populations"]
    pos = oracle.set_particle_positions(pops, (1e-4, 1e-4, 1e-4), 42)
    vel = oracle.set_particle_velocities(pops, params["particle_temperatures"], 42)
    assert pos.shape == (200, 3)
    assert np.all(((pos / 1e-4) ** 2).sum(1) <= 1.0)
    # mt19937 is re-seeded per species (penning_trap.cc:77): equal-mass species at equal T and
    # count get identical velocity sets.
    np.testing.assert_array_equal(vel[:100], vel[100:])
    np.testing.assert_allclose(vel[:100].mean(0), 0.0, atol=1e-9)
