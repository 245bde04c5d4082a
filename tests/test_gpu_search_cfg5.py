"""The surrogate-guided search on the full hyperparameter grid.

Grid: cutoff {5,6,7,8} x Gaussians {6,10,14,18} x max_l {4,6,8,10} x
polynomial order {2,3,4}, 160 candidates by seed 5, 111 of them valid
(polynomial order 4 is rejected, as the reference rejects it). The
exhaustive grid was fitted on the GPU by tests/golden/make_cfg5_exhaustive.py
(200 training and 50 test structures of the binary alloy, mixed precision),
which wrote tests/golden/cfg5_exhaustive.json: test energy and force RMSEs
and the evaluation cost of every candidate. That table is the judge.

The optimal models follow the reference's rule (convex hull of (cost, energy
RMSE), no vertex beaten in cost, energy and force RMSE at once, energy RMSE
below 30 meV/atom; utils/grid_optimal.py:32-68).

CPU: the table covers the grid; its optimal set; the distribution of the
optimal models a random search with the same number of fits (32 seeded + 24
random) recovers, over 2000 seeds.
GPU: the search itself, with the network trained on the GPU by the library's
own kernels on the 32 seeded fits and picking 24 more (fits replayed from the
table): it must recover more optimal models than random search does on
average and at least as many as 90 % of the random draws, and its held-out
predictions must rank the real energy RMSE of the 79 unfitted candidates
(Spearman above a stated bound). A few candidates are re-fitted live on the
GPU to check that the table matches the current code.
"""

import json
import os

import numpy as np
import pytest

from pypolymlp_amd import capi, search

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "golden", "cfg5_exhaustive.json")
N_INIT, N_PICK = 32, 24

needs_table = pytest.mark.skipif(not os.path.exists(TABLE), reason="cfg5 table not generated")


def _table():
    with open(TABLE) as f:
        return json.load(f)


def _results(table):
    by_key = {c.key: c for c in search.grid(160, seed=5)}
    out = {}
    for r in table["results"]:
        c = by_key[r["key"]]
        if r["ok"]:
            out[c] = search.Result(c, True, r["F"], r["rmse_e"], r["rmse_f"], r["cost"],
                                   r["seconds"])
        else:
            out[c] = search.Result(c, False, error="rejected")
    return out


def _front(results):
    good = [r for r in results if r.ok]
    idx = search.find_optimal([r.cost for r in good], [r.rmse_e for r in good],
                              [r.rmse_f for r in good])
    return {good[i].cand.key for i in idx}


def _random_recalls(res, n_seeds=2000):
    valid = [c for c, r in res.items() if r.ok]
    front = _front(res.values())
    rec = []
    for seed in range(n_seeds):
        rng = np.random.default_rng(seed)
        pick = rng.choice(len(valid), N_INIT + N_PICK, replace=False)
        rec.append(len(_front([res[valid[i]] for i in pick]) & front))
    return np.array(rec), front


def _spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


@needs_table
def test_cfg5_table_covers_grid():
    t = _table()
    keys = [r["key"] for r in t["results"]]
    grid = search.grid(160, seed=5)
    assert keys == [c.key for c in grid]
    ok = [r for r in t["results"] if r["ok"]]
    assert len(ok) == 111
    assert all(not r["ok"] for r, c in zip(t["results"], grid) if c.max_p == 4)
    res = _results(t)
    rec, front = _random_recalls(res, 500)
    assert len(front) >= 3
    print(f"optimal models: {sorted(front)}; random search recovers {rec.mean():.2f} on average",
          flush=True)


@needs_table
@pytest.mark.gpu
def test_cfg5_guided_search_beats_random():
    t = _table()
    res = _results(t)
    cands = search.grid(160, seed=5)
    rec, front = _random_recalls(res)
    print(f"cfg5: {len(front)} optimal models; random search (56 fits) recovers "
          f"mean {rec.mean():.2f}, 90th percentile {np.percentile(rec, 90):.0f}", flush=True)
    rep = search.run_search(cands, evaluate_fn=None, n_init=N_INIT, n_pick=N_PICK, seed=7,
                            cached=res)
    assert len(rep.guided) == N_INIT + N_PICK
    assert set(rep.front_exhaustive) == front
    got = len(set(rep.front_guided) & front)
    # held-out ranking of the real energy RMSE by the network trained on the
    # 32 seeded fits
    seeded = {r.cand for r in rep.guided[:N_INIT]}
    held = [r for c, r in res.items() if r.ok and c not in seeded]
    pred = rep.surrogate.predict(np.array([search.descriptor(r.cand, r.n_features)
                                           for r in held]))
    rho_e = _spearman(pred[:, 0], np.log10([r.rmse_e for r in held]))
    rho_c = _spearman(pred[:, 1], np.log10([r.cost for r in held]))
    print(f"guided search recovers {got} of {len(front)}; held-out Spearman: energy RMSE "
          f"{rho_e:.3f}, cost {rho_c:.3f}; hypervolume guided {rep.hv_guided:.4f} random "
          f"{rep.hv_random:.4f} exhaustive {rep.hv_exhaustive:.4f}", flush=True)
    assert rep.surrogate.history[-1] < rep.surrogate.history[0]
    assert got > rec.mean()
    assert got >= np.percentile(rec, 90)
    assert rho_e > RHO_E_MIN
    assert rho_c > RHO_C_MIN


# held-out Spearman bounds; an FP64 numpy restatement of the same training
# and picking rule on this table gave 0.93-0.97 (energy RMSE) and 0.99 (cost)
# over five seeds, and recovered 7.8 of the 8 optimal models on average
RHO_E_MIN = 0.8
RHO_C_MIN = 0.9


@needs_table
@pytest.mark.gpu
def test_cfg5_table_matches_live_fits():
    """Three candidates of the table re-fitted on the GPU now: energy and force
    RMSEs within 2 % (the mixed fit's quantisation and atomic-order noise)."""
    t = _table()
    res = _results(t)
    train, test = search.search_data(250, 50)
    cs = search.cost_cell(4)
    for key in ("rc5_g6_l4_p2", "rc6_g10_l6_p3", "rc8_g18_l4_p2"):
        c = next(c for c in res if c.key == key)
        assert res[c].ok
        print(f"live fit {key}", flush=True)
        r = search.evaluate(c, train, test, precision=capi.PLM_MIXED, cost_structures=cs)
        assert r.rmse_e == pytest.approx(res[c].rmse_e, rel=2e-2)
        assert r.rmse_f == pytest.approx(res[c].rmse_f, rel=2e-2)
