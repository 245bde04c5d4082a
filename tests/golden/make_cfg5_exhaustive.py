"""Writes tests/golden/cfg5_exhaustive.json: the exhaustive hyperparameter grid
of the surrogate-guided search, every valid candidate fitted on the GPU.

Grid: cutoff {5,6,7,8} x Gaussians {6,10,14,18} x max_l {4,6,8,10} x
polynomial order {2,3,4}, 160 of the 192 points by seed 5 (search.grid);
polynomial order 4 is rejected like the reference rejects it. Data: 200
training and 50 test structures of the seeded binary-alloy set
(search.search_data). Each candidate: mixed-precision ridge fit over
alpha {1e-4, 1e-3, 1e-2} chosen by test RMSE, FP64 evaluation of the test
structures (energy RMSE in meV/atom, force RMSE in eV/A) and the evaluation
cost in ms per atom per step on 50 copies of a 256-atom cell.

Needs a GPU (about 5 minutes):
    python tests/golden/make_cfg5_exhaustive.py [out.json]
"""

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from pypolymlp_amd import capi, search  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg5_exhaustive.json")


def main(out=OUT):
    train, test = search.search_data(250, 50)
    cs = search.cost_cell()
    rows = []
    for c in search.grid(160, seed=5):
        if search.feature_count(c) is None:
            rows.append({"key": c.key, "ok": False})
            continue
        r = search.evaluate(c, train, test, precision=capi.PLM_MIXED, cost_structures=cs)
        rows.append({"key": c.key, "ok": True, "F": r.n_features, "rmse_e": r.rmse_e,
                     "rmse_f": r.rmse_f, "cost": r.cost, "seconds": r.seconds})
        print(f"{c.key}: F={r.n_features} rmse_e={r.rmse_e:.4f} rmse_f={r.rmse_f:.5f} "
              f"cost={r.cost:.6f} ({r.seconds:.1f} s)", flush=True)
    with open(out, "w") as f:
        json.dump({"grid": "search.grid(160, seed=5)",
                   "data": "search.search_data(250, 50): 200 train / 50 test",
                   "precision": "mixed", "alphas": [1e-4, 1e-3, 1e-2],
                   "units": {"rmse_e": "meV/atom", "rmse_f": "eV/A", "cost": "ms/atom/step",
                             "seconds": "s (fit + errors + cost)"},
                   "results": rows}, f, indent=0)
    print(f"{sum(r['ok'] for r in rows)} fitted -> {out}", flush=True)


if __name__ == "__main__":
    main(*sys.argv[1:])
