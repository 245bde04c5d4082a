"""Energy-only datasets (include_force False) in the fused X^T X path.

The reference builds only energy rows for such a dataset (Features,
src/pypolymlp/mlp_dev/core/features.py:94-133; PyModel::set_index,
py_model.cpp:58-106), and counts only those rows in total_n_data
(data_sequential.py:155). The device batch always holds the force and
stress rows, so they must carry weight 0 and drop out of the row count.

CPU: row_weights / reference_n_rows closed forms, and the weighted normal
equations of the oracle's X through the batch row weights against the
reference-layout weighting of calc_xy (rows of energy-only datasets
dropped).
GPU: calc_xtx_xty on [dataset with forces, energy-only dataset] against
X^T X / X^T y of calc_xy on the same data, 1e-10 normwise (FP64).
"""

import sys

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi, datasets, mlp_dev


def _two_datasets():
    full = datasets.config1_dataset(10, seed=21)
    a = full.subset(range(6), name="with_forces")
    b = full.subset(range(6, 10), name="energy_only")
    b.include_force = False
    b.include_stress = False
    return a, b


def test_row_weights_energy_only():
    _, b = _two_datasets()
    w, y = mlp_dev.row_weights(b, min_e=-5.0)
    ne = len(b.energies)
    assert len(w) == ne + 6 * ne + 3 * int(b.total_n_atoms.sum())
    assert np.all(w[:ne] > 0)
    assert np.all(w[ne:] == 0.0) and np.all(y[ne:] == 0.0)
    np.testing.assert_allclose(y[:ne], w[:ne] * b.energies, rtol=1e-15)
    assert mlp_dev.reference_n_rows(b) == ne
    assert mlp_dev.reference_n_rows(b, 2) == 2


def test_row_weights_with_forces_no_stress():
    a, _ = _two_datasets()
    a.include_stress = False
    w, y = mlp_dev.row_weights(a, min_e=-5.0)
    ne = len(a.energies)
    assert np.all(w[ne:7 * ne] == 0.0) and np.all(y[ne:7 * ne] == 0.0)
    f = a.forces_flat
    np.testing.assert_allclose(w[7 * ne:], np.minimum(1.0, 1.0 / np.maximum(np.abs(f), 1e-12)))
    assert mlp_dev.reference_n_rows(a) == 7 * ne + 3 * int(a.total_n_atoms.sum())


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
def test_batch_weights_reproduce_reference_layout(tmp_path):
    """The oracle's X of each dataset in batch layout, weighted by row_weights,
    gives the same X^T X and X^T y as the reference layout in which the
    energy-only dataset has energy rows only."""
    from tests.helpers import write_model_yaml
    spec = datasets.config1_model()
    R = ref.RefModel(write_model_yaml(spec, str(tmp_path / "m.yaml"), np.zeros(2277)))
    a, b = _two_datasets()
    min_e = mlp_dev.get_min_energy([a, b])
    xtx = np.zeros((2277, 2277))
    xty = np.zeros(2277)
    for d in (a, b):
        sts = d.structures
        X = R.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                      [s.types for s in sts], force=True)
        w, y = mlp_dev.row_weights(d, min_e)
        Xw = X * w[:, None]
        xtx += Xw.T @ Xw
        xty += Xw.T @ y
    # reference layout: energy-only rows for b
    sts = b.structures
    Xe = R.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts], force=False)
    we = mlp_dev._weight_energy(b.energies, b.total_n_atoms, min_e=min_e)
    sts = a.structures
    Xa = R.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts], force=True)
    wa, ya = mlp_dev.row_weights(a, min_e)
    Xr = np.vstack([Xa * wa[:, None], Xe * we[:, None]])
    yr = np.concatenate([ya, we * b.energies])
    assert np.linalg.norm(xtx - Xr.T @ Xr) <= 1e-12 * np.linalg.norm(Xr.T @ Xr)
    assert np.linalg.norm(xty - Xr.T @ yr) <= 1e-12 * np.linalg.norm(Xr.T @ yr)
    assert Xr.shape[0] == mlp_dev.reference_n_rows(a) + mlp_dev.reference_n_rows(b)


@pytest.mark.gpu
def test_calc_xtx_xty_matches_calc_xy_with_energy_only_dataset():
    print("energy-only: cfg1 model, 6 + 4 structures, FP64", flush=True)
    sys.stdout.flush()
    spec = datasets.config1_model()
    model = capi.Model(spec.as_dict())
    a, b = _two_datasets()
    ne = mlp_dev.calc_xtx_xty(model, [a, b], precision=capi.PLM_FP64)
    xy = mlp_dev.calc_xy(model, [a, b], scales=ne.scales.copy(), min_energy=ne.min_energy)
    # columns calc_xtx_xty dropped for a zero scale (their scale is now 1)
    xy.x[:, np.diag(ne.xtx) == 0.0] = 0.0
    xtx = xy.x.T @ xy.x
    xty = xy.x.T @ xy.y
    assert xy.x.shape[0] == ne.total_n_data
    assert np.linalg.norm(ne.xtx - xtx) <= 1e-10 * np.linalg.norm(xtx)
    assert np.linalg.norm(ne.xty - xty) <= 1e-10 * np.linalg.norm(xty)
    assert ne.y_sq_norm == pytest.approx(float(xy.y @ xy.y), rel=1e-12)
