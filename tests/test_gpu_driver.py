"""The native driver (am_simulate, and the am_simulate executable) on the shipped params.json: the
energy CSV has the reference's header and number format (std::ostream defaults, 6 significant
digits, antimatter-kernel.cc:128,146-150,189-192), and its rows 0-4 read back as the notebook's
stored values (simulation_analysis.ipynb:161-172)."""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import antimatter_amd as am
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

HEADER = ["step", "kinetic_energy_J", "trap_potential_energy_J", "potential_energy_J", "total_energy_J",
          "num_active_particles", "fractional_drift", "absolute_drift_J"]
PARAMS = os.path.join(GOLDEN, "params_config1.json")


def read_csv(path):
    with open(path) as fh:
        rows = list(csv.reader(fh))
    return rows[0], rows[1:]


def test_energy_csv_matches_notebook(tmp_path):
    golden = json.load(open(os.path.join(GOLDEN, "config1_notebook.json")))
    out = str(tmp_path / "energy_analysis_data.csv")
    rows = am.simulate(PARAMS, csv_path=out)
    header, body = read_csv(out)
    assert header == HEADER
    assert len(body) == len(rows) == 101
    for g in golden["rows"]:
        r = body[g["step"]]
        assert int(r[0]) == g["step"] and int(r[5]) == g["num_active"]
        # Within one unit of the 6th significant digit: the reference's neighbour-order-dependent sums
        # move the unrounded values by ~1e-9 relative, which can flip a last digit that sits on a
        # rounding boundary (step 3's absolute drift: -1.537475169e-20 here).
        for col, key in ((1, "kinetic"), (2, "trap_potential"), (3, "energy_potential"), (4, "total"),
                         (7, "absolute_drift")):
            ulp6 = 10.0 ** (np.floor(np.log10(max(abs(g[key]), 1e-300))) - 5)
            assert abs(float(r[col]) - g[key]) <= 1.01 * ulp6, (g["step"], key, r[col])
            assert len(r[col].split("e")[0].replace("-", "").replace(".", "")) <= 6  # 6 significant digits
        assert abs(round(float(r[6]), 6) - g["fractional_drift"]) <= 1.01e-6
    assert int(body[-1][5]) == golden["active_end"]
    assert round(float(body[-1][6]), 6) == golden["final_fractional_drift"]
    # the CSV holds the returned rows at 6 significant digits
    np.testing.assert_allclose([float(x) for x in body[50][1:5]], rows[50, 1:5], rtol=5e-6)


def test_executable_writes_the_same_csv(tmp_path):
    exe = os.path.join(ROOT, "antimatter-kernel_amd", "am_simulate")
    ref = str(tmp_path / "lib.csv")
    am.simulate(PARAMS, csv_path=ref)
    outdir = tmp_path / "out"
    res = subprocess.run([exe, "--params", PARAMS, "--output-dir", str(outdir)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Simulation completed successfully!" in res.stdout
    assert "step 100  KE=" in res.stdout and "of trap depth)" in res.stdout
    assert open(outdir / "energy_analysis_data.csv").read() == open(ref).read()
    # an invalid configuration stops before any device work, with the reason
    bad = tmp_path / "bad.json"
    cfg = json.load(open(PARAMS))
    cfg["coulomb_computation_method"] = "barnes_hut"
    bad.write_text(json.dumps(cfg))
    res = subprocess.run([exe, "--params", str(bad), "--output-dir", str(outdir)], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 1 and "barnes_hut" in res.stderr
