"""Writes tests/golden/grid_ti_summary.json from the reference's grid-Ti test
data (tests/files/grid-Ti/polymlp-*/polymlp_{error,cost}.yaml): per model
directory its name, single-core and OpenMP cost (ms/atom/step), and the
energy / force RMSEs of the first dataset whose name contains
"test_vasprun_low" (the key of tests/test_utils/test_grid_optimal.py:14).
Directories without that dataset or without a cost file are skipped, as the
reference's parser skips them (utils/grid_optimal.py:120-143).

Run in the build container only (it reads /root/reference):
    python tests/golden/make_grid_ti.py
"""

import glob
import json
import os

import yaml

SRC = "/root/reference/tests/files/grid-Ti"
KEY = "test_vasprun_low"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grid_ti_summary.json")


def main():
    rows = []
    for d in sorted(glob.glob(os.path.join(SRC, "polymlp-*"))):
        fe, fc = os.path.join(d, "polymlp_error.yaml"), os.path.join(d, "polymlp_cost.yaml")
        if not (os.path.exists(fe) and os.path.exists(fc)):
            continue
        err = yaml.safe_load(open(fe))
        values = err.get("prediction_errors", err.get("prediction_errors_test"))
        if values is None:
            continue
        match = next((v for v in values if KEY in v["dataset"]), None)
        if match is None:
            continue
        cost = yaml.safe_load(open(fc))["costs"]
        rows.append({"name": os.path.basename(d), "cost_single": cost["single_core"],
                     "cost_openmp": cost["openmp"], "rmse_energy": match["rmse_energy"],
                     "rmse_force": match["rmse_force"]})
    with open(OUT, "w") as f:
        json.dump({"source": "tests/files/grid-Ti (reference test data), key " + KEY,
                   "models": rows}, f, indent=0)
    print(f"{len(rows)} models -> {OUT}")


if __name__ == "__main__":
    main()
