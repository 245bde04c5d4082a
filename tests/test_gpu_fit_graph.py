"""The fit step captured as a HIP graph (plm_normal_graph_*) against eager
accumulation.

The binary-alloy model (linear, comb2 and comb3 columns: every kernel class
of the X build, the polynomial GEMMs included) on 12 structures, split into
several sub-batches by a small X budget. Creation leaves the accumulator at
zero; each replay adds one pass. Replays run with other eager accumulations
and synchronisations between them, and every replay is checked against the
eager sums: FP64 to 1e-10 and mixed to 1e-6 (relative, normwise). The X build
adds per-pair terms with atomics (in the mixed form FP32 adds into the factor
derivatives), so the last bits of X, and in the mixed form an occasional
quantisation digit, depend on their order: two eager mixed accumulations of
this batch differ by up to 5e-8 in X^T X and 1.1e-7 in X^T y (measured), FP64
ones by 2e-16. A replay that missed a sub-batch or read a stale buffer would be
off by 1e-3 or more.
"""

import os
import sys

import numpy as np
import pytest

from pypolymlp_amd import capi, datasets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [(capi.PLM_FP64, 1e-10), (capi.PLM_MIXED, 1e-6)])
def test_fit_graph_replays_match_eager(precision, tol, monkeypatch):
    print(f"fit graph: binary alloy model, 12 structures, precision {precision}", flush=True)
    sys.stdout.flush()
    monkeypatch.setenv("PLM_X_BUDGET_MB", "2")
    ds = datasets.config2_dataset(12, seed=31, targets=False)
    model = capi.Model(datasets.config2_model().as_dict())
    sts = ds.structures
    batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                       [s.types for s in sts])
    rng = np.random.default_rng(7)
    w = rng.uniform(0.1, 1.0, batch.n_rows)
    y = rng.normal(size=batch.n_rows)
    eager = capi.Normal(model)
    eager.accumulate(batch, w, y, precision)
    x1, v1, s1, q1, y1, n1 = eager.result()
    acc = capi.Normal(model)
    graph = capi.NormalGraph(acc, batch, w, y, precision)
    x0, _, _, _, y0, n0 = acc.result()
    assert not x0.any() and y0 == 0.0 and n0 == 0

    def close(a, b):
        return np.linalg.norm(a - b) <= tol * np.linalg.norm(b)

    other = capi.Normal(model)
    for k in (1, 2, 3):
        graph.run(1)
        xk, vk, sk, qk, yk, nk = acc.result()
        assert nk == k * n1
        assert close(xk, k * x1) and close(vk, k * v1)
        assert close(sk, k * s1) and close(qk, k * q1)
        assert yk == pytest.approx(k * y1, rel=tol)
        other.accumulate(batch, w, y, precision)  # eager work between replays
    graph.run(2)
    assert close(acc.result()[0], 5 * x1)
