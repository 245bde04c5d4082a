"""Synthetic datasets and models of the benchmark configurations.

Every structure set is generated from a fixed seed with
``numpy.random.default_rng(seed)``. Cartesian positions are ``axis @ frac``
without rewrapping (as ``src/pypolymlp/mlp_dev/core/features.py:16`` passes
them). Targets come from a (multi-species) Morse pair potential

    E = 1/2 sum_{i != j, r < rc} phi(r),  phi = D [(1 - exp(-a (r - r0)))^2 - 1]

with forces -dE/dr and stress in the reference convention
s_ab = -sum_{unique pairs} phi'(r) d_a d_b / r, ordered (xx, yy, zz, xy, yz, zx)
in eV/cell (``src/cxx/src/compute/polymlp_eval.cpp:293-335``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from .params import ModelSpec
from .structure import PolymlpStructure


@dataclass
class Dataset:
    structures: List[PolymlpStructure]
    energies: np.ndarray
    forces: List[np.ndarray]  # each (3, N)
    stresses: np.ndarray  # (n_st, 6)
    name: str = "synthetic"
    weight: float = 1.0
    include_force: bool = True
    include_stress: bool = True
    exist_force: bool = True   # the source held forces (DatasetDFT.exist_force)
    exist_stress: bool = True  # the source held stresses; zeros otherwise

    @property
    def total_n_atoms(self) -> np.ndarray:
        return np.array([s.n_total for s in self.structures])

    @property
    def forces_flat(self) -> np.ndarray:
        """Forces in the reference's flattened order (3k + a per structure)."""
        return np.concatenate([f.T.reshape(-1) for f in self.forces])

    @property
    def stresses_flat(self) -> np.ndarray:
        return self.stresses.reshape(-1)

    def subset(self, ids, name=None) -> "Dataset":
        """Structures ``ids`` (an index sequence or a slice) with their targets."""
        ids = np.arange(len(self.structures))[ids] if isinstance(ids, slice) else np.asarray(ids)
        return Dataset(structures=[self.structures[i] for i in ids], energies=self.energies[ids],
                       forces=[self.forces[i] for i in ids], stresses=self.stresses[ids],
                       name=self.name if name is None else name, weight=self.weight,
                       include_force=self.include_force, include_stress=self.include_stress,
                       exist_force=self.exist_force, exist_stress=self.exist_stress)

    def split(self, train_ratio: float = 0.9):
        """(train, test) by split_ids_train_test (DatasetDFT.split,
        src/pypolymlp/core/dataset_utils.py:351-387; names as Dataset.split_dft)."""
        tr, te = split_ids_train_test(len(self.energies), train_ratio=train_ratio)
        train = self.subset(tr, "Train_" + self.name) if len(tr) else None
        test = self.subset(te, "Test_" + self.name) if len(te) else None
        return train, test

    def apply_atomic_energy(self, atomic_energy):
        """Subtract sum_t n_t e_t from each energy (DatasetDFT.apply_atomic_energy)."""
        atomic_energy = np.asarray(atomic_energy, dtype=float)
        self.energies = np.array([e - float(np.dot(s.n_atoms, atomic_energy[:len(s.n_atoms)]))
                                  for e, s in zip(self.energies, self.structures)])
        return self

    def sort_by_atoms(self):
        """Stable sort by atom count (Dataset.sort_dft, dataset_utils.py:337-349)."""
        order = np.argsort([s.n_total for s in self.structures], kind="stable")
        self.structures = [self.structures[i] for i in order]
        self.energies = self.energies[order]
        self.forces = [self.forces[i] for i in order]
        self.stresses = self.stresses[order]
        return self


def split_ids_train_test(n_data: int, train_ratio: float = 0.9):
    """Test ids evenly spread over [0, n_data) by rounding a linspace, the rest
    train (src/pypolymlp/core/utils.py:18-27)."""
    n_train = round(n_data * train_ratio)
    n_test = n_data - n_train
    test_ids = np.round(np.linspace(0, n_data - 1, num=n_test)).astype(int)
    train = np.ones(n_data, dtype=bool)
    train[test_ids] = False
    return np.where(train)[0], test_ids


@dataclass
class MorseTable:
    D: np.ndarray
    alpha: np.ndarray
    r0: np.ndarray


def _images(axis: np.ndarray, cutoff: float) -> np.ndarray:
    inv = np.linalg.inv(axis)
    # number of images along each lattice vector from the reciprocal heights
    heights = 1.0 / np.linalg.norm(inv, axis=1)
    n = np.ceil(cutoff / heights).astype(int) + 1
    g = np.mgrid[-n[0]:n[0] + 1, -n[1]:n[1] + 1, -n[2]:n[2] + 1].reshape(3, -1).T
    return g @ axis.T


def morse_targets(st: PolymlpStructure, table: MorseTable, cutoff: float):
    """(energy, forces (3, N), stress (6)) of the Morse potential."""
    pos = (st.axis @ st.positions).T
    types = np.asarray(st.types)
    n = len(types)
    trans = _images(st.axis, cutoff)
    e = 0.0
    f = np.zeros((n, 3))
    s = np.zeros(6)
    for i in range(n):
        d = pos[None, :, :] - pos[i][None, None, :] + trans[:, None, :]  # (T, N, 3)
        d = d.reshape(-1, 3)
        r = np.linalg.norm(d, axis=1)
        j = np.tile(np.arange(n), len(trans))
        mask = (r < cutoff) & (r > 1e-10)
        d, r, j = d[mask], r[mask], j[mask]
        D = table.D[types[i], types[j]]
        a = table.alpha[types[i], types[j]]
        r0 = table.r0[types[i], types[j]]
        ex = np.exp(-a * (r - r0))
        phi = D * ((1.0 - ex) ** 2 - 1.0)
        dphi = 2.0 * D * (1.0 - ex) * a * ex
        e += 0.5 * phi.sum()
        # d = r_j - r_i: dE/dr_i = -sum phi'(r) d / r (pair counted from both ends)
        f[i] += (dphi / r) @ d
        g = 0.5 * dphi / r
        s[0] -= np.sum(g * d[:, 0] * d[:, 0])
        s[1] -= np.sum(g * d[:, 1] * d[:, 1])
        s[2] -= np.sum(g * d[:, 2] * d[:, 2])
        s[3] -= np.sum(g * d[:, 0] * d[:, 1])
        s[4] -= np.sum(g * d[:, 1] * d[:, 2])
        s[5] -= np.sum(g * d[:, 2] * d[:, 0])
    return e, f.T, s


def _bcc_supercell(a: float, reps):
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    return _supercell(np.eye(3) * a, basis, reps)


def _fcc_supercell(a: float, reps):
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    return _supercell(np.eye(3) * a, basis, reps)


def _supercell(axis0, basis, reps):
    reps = np.asarray(reps)
    grid = np.mgrid[0:reps[0], 0:reps[1], 0:reps[2]].reshape(3, -1).T
    frac = (grid[:, None, :] + basis[None, :, :]).reshape(-1, 3) / reps[None, :]
    axis = axis0 @ np.diag(reps)
    return axis, frac.T


def _make(axis, frac, types, elements, rng, sigma, scale=1.0):
    axis = axis * scale
    cart = axis @ frac + rng.normal(0.0, sigma, size=frac.shape)
    pos = np.linalg.solve(axis, cart)
    n_type = len(elements)
    counts = [int(np.sum(np.asarray(types) == t)) for t in range(n_type)]
    return PolymlpStructure(axis=axis, positions=pos, n_atoms=counts,
                            elements=[elements[t] for t in types], types=list(map(int, types)))


def _finish(structures, table, cutoff, name):
    es, fs, ss = [], [], []
    for st in structures:
        e, f, s = morse_targets(st, table, cutoff)
        es.append(e), fs.append(f), ss.append(s)
    return Dataset(structures=structures, energies=np.array(es), forces=fs,
                   stresses=np.array(ss), name=name)


def config1_model() -> ModelSpec:
    return ModelSpec(n_type=1, cutoff=6.0, model_type=4, max_p=2, gtinv_order=2,
                     gtinv_maxl=(4,), gaussian_params1=(1.0, 1.0, 1),
                     gaussian_params2=(0.0, 5.0, 10), elements=["A"])


def config1_dataset(n_st: int = 200, seed: int = 1) -> Dataset:
    """200 x 54-atom bcc 3x3x3 (a = 3.30), strain +-3 %, sigma = 0.08 A."""
    rng = np.random.default_rng(seed)
    axis, frac = _bcc_supercell(3.30, (3, 3, 3))
    table = MorseTable(D=np.array([[0.5]]), alpha=np.array([[1.5]]), r0=np.array([[2.86]]))
    sts = []
    for _ in range(n_st):
        eps = rng.uniform(-0.03, 0.03)
        sts.append(_make(axis, frac, [0] * frac.shape[1], ["A"], rng, 0.08, 1.0 + eps))
    return _finish(sts, table, 6.0, "cfg1")


def config2_model() -> ModelSpec:
    return ModelSpec(n_type=2, cutoff=7.0, model_type=3, max_p=3, gtinv_order=3,
                     gtinv_maxl=(8, 4), gaussian_params1=(1.0, 1.0, 1),
                     gaussian_params2=(0.0, 6.0, 14), elements=["A", "B"])


def config2_dataset(n_st: int = 1000, seed: int = 2, targets: bool = True) -> Dataset:
    """1000 x 128-atom fcc 4x4x2 (a = 3.9) binary, volume +-8 %, sigma = 0.1 A."""
    rng = np.random.default_rng(seed)
    axis, frac = _fcc_supercell(3.9, (4, 4, 2))
    table = MorseTable(D=np.array([[0.4, 0.6], [0.6, 0.5]]), alpha=np.full((2, 2), 1.4),
                       r0=np.full((2, 2), 2.75))
    sts = []
    for _ in range(n_st):
        c = rng.uniform(0.0, 1.0)
        types = (rng.uniform(size=frac.shape[1]) < c).astype(int)
        v = rng.uniform(-0.08, 0.08)
        sts.append(_make(axis, frac, types, ["A", "B"], rng, 0.1, (1.0 + v) ** (1.0 / 3.0)))
    if not targets:
        n = len(sts)
        return Dataset(structures=sts, energies=np.zeros(n),
                       forces=[np.zeros((3, s.n_total)) for s in sts], stresses=np.zeros((n, 6)))
    return _finish(sts, table, 7.0, "cfg2")


def config3_model() -> ModelSpec:
    return ModelSpec(n_type=3, cutoff=6.0, model_type=3, max_p=2, gtinv_order=4,
                     gtinv_maxl=(8, 4, 2), gaussian_params1=(1.0, 1.0, 1),
                     gaussian_params2=(0.0, 5.0, 12), elements=["A", "B", "C"])


def config3_table(seed: int = 3) -> MorseTable:
    rng = np.random.default_rng(seed + 1000)
    D = rng.uniform(0.3, 0.7, size=(3, 3))
    r0 = rng.uniform(2.5, 3.0, size=(3, 3))
    D = 0.5 * (D + D.T)
    r0 = 0.5 * (r0 + r0.T)
    return MorseTable(D=D, alpha=np.full((3, 3), 1.5), r0=r0)


def config3_dataset(n_st: int = 5000, seed: int = 3, targets: bool = True) -> Dataset:
    """Ternary structures with N uniform in [8, 432] on random cubic / tetragonal /
    hexagonal lattices (density +-15 %), random species, sigma = 0.12 A. Sites
    beyond N of the smallest near-cubic supercell holding N are left empty."""
    rng = np.random.default_rng(seed)
    sts = []
    d_nn = 2.7
    for _ in range(n_st):
        n = int(rng.integers(8, 433))
        kind = int(rng.integers(0, 3))
        dens = 1.0 + rng.uniform(-0.15, 0.15)
        if kind == 0:
            axis0, basis = np.eye(3) * d_nn, np.zeros((1, 3))
        elif kind == 1:
            c = rng.uniform(1.05, 1.3)
            axis0, basis = np.diag([d_nn, d_nn, d_nn * c]), np.zeros((1, 3))
        else:
            a = d_nn
            c = a * rng.uniform(1.5, 1.7)
            axis0 = np.array([[a, -0.5 * a, 0.0], [0.0, np.sqrt(3) / 2 * a, 0.0], [0.0, 0.0, c]])
            basis = np.array([[0, 0, 0], [1.0 / 3.0, 2.0 / 3.0, 0.5]])
        nb = len(basis)
        k = int(np.ceil((n / nb) ** (1.0 / 3.0)))
        reps = [k, k, k]
        while np.prod(reps) * nb - nb * reps[0] * reps[1] >= n and reps[2] > 1:
            reps[2] -= 1
        axis, frac = _supercell(axis0, basis, reps)
        sel = np.sort(rng.choice(frac.shape[1], size=n, replace=False))
        frac = frac[:, sel]
        types = rng.integers(0, 3, size=n)
        scale = dens ** (-1.0 / 3.0)
        sts.append(_make(axis, frac, types, ["A", "B", "C"], rng, 0.12, scale))
    if not targets:
        m = len(sts)
        return Dataset(structures=sts, energies=np.zeros(m),
                       forces=[np.zeros((3, s.n_total)) for s in sts], stresses=np.zeros((m, 6)))
    return _finish(sts, config3_table(seed), 6.0, "cfg3")


def config4_structure(seed: int = 4, reps=(20, 20, 40)) -> PolymlpStructure:
    """32,000-atom bcc supercell (20 x 20 x 40, a = 3.30), sigma = 0.1 A."""
    rng = np.random.default_rng(seed)
    axis, frac = _bcc_supercell(3.30, reps)
    return _make(axis, frac, [0] * frac.shape[1], ["A"], rng, 0.1)
