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


def read_vtu(path):
    import xml.etree.ElementTree as ET
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
    arrays = {a.get("Name"): a for a in piece.iter("DataArray")}
    npts = int(piece.get("NumberOfPoints"))
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=np.float64).reshape(npts, 3)
    sid = np.array(arrays["species_id"].text.split(), dtype=np.int64)
    diam = np.array(arrays["diameter_"].text.split(), dtype=np.float64)
    conn = np.array(arrays["connectivity"].text.split(), dtype=np.int64)
    types = np.array(arrays["types"].text.split(), dtype=np.int64)
    return pts, sid, diam, conn, types


def test_paraview_export(tmp_path):
    """The agents of every 10th step as VTK XML (the reference's BioDynaMo export, bdm.toml:6-13, with
    species_id and the 1e-6 m diameter), and a .pvd collection of them."""
    import xml.etree.ElementTree as ET
    from antimatter_amd import params as P
    rows = am.simulate(PARAMS, export_dir=str(tmp_path), export_interval=10)
    cfg = P.load_params(PARAMS)
    coll = ET.parse(tmp_path / "ChargedAntiparticle.pvd").getroot().find("Collection")
    sets = coll.findall("DataSet")
    assert [d.get("file") for d in sets] == [f"ChargedAntiparticle-{k}.vtu" for k in range(0, 101, 10)]
    np.testing.assert_allclose([float(d.get("timestep")) for d in sets], np.arange(0, 101, 10) * cfg.dt, rtol=1e-15)
    # step 0: the reference's initial conditions, 100 positrons (species 1) and 100 electrons (3)
    pos0, _, sp0 = am.initial_conditions({"positron": 100, "electron": 100}, {"positron": 1000.0, "electron": 1000.0},
                                         (1e-4, 1e-4, 1e-4), 42)
    pts, sid, diam, conn, types = read_vtu(tmp_path / "ChargedAntiparticle-0.vtu")
    np.testing.assert_array_equal(pts, pos0)  # 17 significant digits round-trip exactly
    np.testing.assert_array_equal(sid, sp0)
    assert np.all(diam == 1e-6)
    np.testing.assert_array_equal(conn, np.arange(200))
    assert np.all(types == 1)  # VTK_VERTEX
    # later steps hold the agents still in the simulation: the logged active counts, pairs removed
    for k in range(10, 101, 10):
        pts, sid, _, _, _ = read_vtu(tmp_path / f"ChargedAntiparticle-{k}.vtu")
        assert len(pts) == int(rows[k, 5]) == len(sid)
        assert (sid == 1).sum() == (sid == 3).sum()  # annihilation removes e+ e- pairs
        assert np.all(np.isfinite(pts)) and np.all(np.abs(pts) < 1e-2)
    assert len(read_vtu(tmp_path / "ChargedAntiparticle-100.vtu")[0]) == 170
    # the executable's --export-interval writes the same files into <output-dir>/visualization
    exe = os.path.join(ROOT, "antimatter-kernel_amd", "am_simulate")
    outdir = tmp_path / "out"
    res = subprocess.run([exe, "--params", PARAMS, "--output-dir", str(outdir), "--export-interval", "50"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    for k in (0, 50, 100):
        name = f"ChargedAntiparticle-{k}.vtu"
        assert open(outdir / "visualization" / name).read() == open(tmp_path / name).read()
