"""Writes cfg3_fit_subset.npz: 80 structures of the ternary batched-uneven
workload (config3_dataset(500, seed=3)): the 30 largest (400+ atoms) and 50
drawn from the rest with seed 0, in that order, with their Morse targets.
Positions and forces are stored in float32 (the fixture feeds a comparison of
two fits on the same data, not a parity check of the targets)."""
import os

import numpy as np

from pypolymlp_amd import datasets

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ds = datasets.config3_dataset(500, seed=3)
    n = np.array([s.n_total for s in ds.structures])
    order = np.argsort(-n, kind="stable")
    big = order[:30]
    rest = np.random.default_rng(0).choice(order[30:], 50, replace=False)
    pick = np.concatenate([big, rest])
    sts = [ds.structures[i] for i in pick]
    np.savez_compressed(
        os.path.join(HERE, "cfg3_fit_subset.npz"),
        axes=np.array([s.axis for s in sts]),
        pos=np.concatenate([s.positions.T for s in sts]).astype(np.float32),
        n=np.array([s.n_total for s in sts]),
        types=np.concatenate([np.asarray(s.types) for s in sts]).astype(np.int8),
        energies=ds.energies[pick],
        forces=np.concatenate([ds.forces[i].T for i in pick]).astype(np.float32),
        stresses=ds.stresses[pick])


if __name__ == "__main__":
    main()
