"""Writes cfg3_large.npz: oracle X of the largest structure of the seeded
40-structure ternary set (429 atoms, 1294 rows x 25,428 columns), kept as
seeded projections u^T X and X v and the energy row.

The oracle (the reference's own C++ built under oracle/_ref) takes about
five minutes for this structure, too long for a test; the fixture stores the
projections so a test can check a device X of that size in seconds.

    python tests/golden/make_cfg3_large.py
"""

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pypolymlp_amd import datasets  # noqa: E402
from tests.helpers import write_model_yaml  # noqa: E402

N_ST, SEED_U, SEED_V = 40, 31, 32


def largest_structure():
    ds = datasets.config3_dataset(N_ST, targets=False)
    return max(ds.structures, key=lambda s: s.n_total)


def projections(X):
    u = np.random.default_rng(SEED_U).normal(size=X.shape[0])
    v = np.random.default_rng(SEED_V).normal(size=X.shape[1])
    return u @ X, X @ v


def main():
    st = largest_structure()
    spec = datasets.config3_model()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "model.yaml")
        write_model_yaml(spec, path, np.ones(spec_features(spec)))
        X = ref.RefModel(path).x_batch([st.axis], [st.positions_cartesian], [st.types], force=True)
    p, q = projections(X)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg3_large.npz")
    np.savez_compressed(out, n_atoms=st.n_total, shape=np.array(X.shape), ut_x=p, x_v=q,
                        energy_row=X[0], col_norm=np.linalg.norm(X, axis=0))
    print(out, X.shape)


def spec_features(spec):
    from pypolymlp_amd import capi
    return capi.Model(spec.as_dict()).n_features


if __name__ == "__main__":
    main()
