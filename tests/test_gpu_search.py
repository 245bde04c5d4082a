"""Surrogate-guided search on a small grid: every candidate is fitted on the
GPU, the network is trained on a seeded subset, and its predictions must rank
the held-out candidates (Spearman correlation of log RMSE and of log cost
above 0.5); the guided search's front consists of fitted candidates only."""

import numpy as np
import pytest

from pypolymlp_amd import capi, datasets, search

pytestmark = pytest.mark.gpu


def _spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


def test_search_small_grid():
    ds = datasets.config2_dataset(36, seed=21)
    def sub(i0, i1):
        return datasets.Dataset(structures=ds.structures[i0:i1], energies=ds.energies[i0:i1],
                                forces=ds.forces[i0:i1], stresses=ds.stresses[i0:i1])
    train, test = sub(0, 28), sub(28, 36)
    cands = [search.Candidate(rc, g, l, p) for rc in (5.0, 6.0, 7.0) for g in (6, 10)
             for l in (4, 6) for p in (2, 3, 4)]
    cs = search.cost_cell(4)
    rep = search.run_search(cands, lambda c: search.evaluate(c, train, test,
                                                             precision=capi.PLM_FP64,
                                                             cost_structures=cs),
                            n_init=12, n_pick=6, surrogate_steps=2000)
    ok = [r for r in rep.exhaustive if r.ok]
    assert len(ok) == 24  # max_p = 4 rejected like the reference
    assert all(not r.ok for r in rep.exhaustive if r.cand.max_p == 4)
    assert rep.surrogate.history[-1] < rep.surrogate.history[0]
    fitted = {r.cand.key for r in rep.guided}
    assert set(rep.front_guided) <= fitted
    held = [r for r in ok if r.cand.key not in {x.cand.key for x in rep.guided[:12]}]
    pred = rep.surrogate.predict(np.array([search.descriptor(r.cand, r.n_features) for r in held]))
    rho_e = _spearman(pred[:, 0], np.log10([r.rmse_e for r in held]))
    rho_c = _spearman(pred[:, 1], np.log10([r.cost for r in held]))
    print(f"held-out Spearman: log rmse {rho_e:.2f}, log cost {rho_c:.2f}; "
          f"hypervolume exhaustive {rep.hv_exhaustive:.3f} guided {rep.hv_guided:.3f} "
          f"random {rep.hv_random:.3f}", flush=True)
    assert rho_c > 0.5
    assert rho_e > 0.5
