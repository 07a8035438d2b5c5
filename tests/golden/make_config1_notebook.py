"""Writes config1_notebook.json from the stored outputs of the reference notebook.

The notebook (simulation_analysis.ipynb) is read as JSON data; nothing in it is executed.
It holds the only published results of the reference: one run of the shipped params.json
(100 e+ + 100 e-, 1000 K, B0 = 1 T, U0 = 10 V, cutoff direct r_c = 1e-5 m, annihilation
r_a = 5e-6 m at p = 1, 100 steps).  The energy CSV it reads was written with the default
6 significant digits (src/antimatter-kernel.cc:146-150,189-192).

Usage: python tests/golden/make_config1_notebook.py /path/to/reference
"""
import json
import os
import re
import sys


def main(ref_root: str) -> None:
    nb = json.load(open(os.path.join(ref_root, "simulation_analysis.ipynb")))
    texts = []
    for cell in nb["cells"]:
        if cell["cell_type"] != "code":
            continue
        for out in cell.get("outputs", []):
            if "text" in out:
                texts.append("".join(out["text"]))
            elif "data" in out and "text/plain" in out["data"]:
                texts.append("".join(out["data"]["text/plain"]))
    blob = "\n".join(texts)
    # Rows 0-4 of the energy table (simulation_analysis.ipynb:161-165).
    rows = []
    for m in re.finditer(r"^(\d)\s+(\d)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\d+)\s+\\?$", blob, re.M):
        rows.append(dict(step=int(m.group(2)), kinetic=float(m.group(3)), trap_potential=float(m.group(4)),
                         energy_potential=float(m.group(5)), total=float(m.group(6)),
                         num_active=int(m.group(7))))
    drift = {}
    for m in re.finditer(r"^(\d)\s+(-?\d\.\d+)\s+(-?\d\.\d+e[-+]\d+)\s+\S+\s+\S+\s+\S+\s+\\?$", blob, re.M):
        drift[int(m.group(1))] = dict(fractional_drift=float(m.group(2)), absolute_drift=float(m.group(3)))
    for r in rows:
        r.update(drift.get(r["step"], {}))
    rows = rows[:5]
    # Drift change per 20-step window (simulation_analysis.ipynb:438-442).
    windows = [float(m.group(1)) for m in re.finditer(r"^\d\s+\d+-\d+\s+(-?\d\.\d+)\s+\S+$", blob, re.M)]
    final = float(re.search(r"final fractional drift\s+(-?\d\.\d+)", blob).group(1))
    maxd = float(re.search(r"max \|fractional drift\|\s+(\d\.\d+)", blob).group(1))
    act = re.search(r"Active particles: (\d+) at start, (\d+) at end", blob)
    fixture = dict(
        source="simulation_analysis.ipynb stored outputs (cells 2, 9, 13, 15)",
        params="params.json as shipped",
        csv_significant_digits=6,
        rows=rows,
        drift_change_per_20_steps=windows,
        final_fractional_drift=final,
        max_abs_fractional_drift=maxd,
        active_start=int(act.group(1)),
        active_end=int(act.group(2)),
    )
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config1_notebook.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
