"""Reader of phonopy / phono3py YAML training datasets.

Host-side mirror of ``src/pypolymlp/core/interface_phono3py.py:15-161``
without the phonopy / phono3py packages: the YAML file (plain, ``.xz``,
``.gz`` or ``.bz2``) is read with PyYAML's safe loader and the displaced
supercells are built as the reference builds them
(``src/pypolymlp/core/displacements.py:10-67``): fractional positions of the
perfect supercell plus ``inv(axis) @ d`` for Cartesian displacements ``d``.

Only displacement datasets of the "type 2" form (a displacement and a force
for every atom of every snapshot: ``dataset.displacements`` /
``dataset.forces``, or ``phonon_dataset.*`` with ``use_phonon_dataset``)
are supported; the per-snapshot energies are ``dataset.supercell_energies``.
"""

from __future__ import annotations

import bz2
import gzip
import lzma
from typing import List, Optional, Sequence

import numpy as np
import yaml

from .datasets import Dataset
from .structure import PolymlpStructure

_Loader = getattr(yaml, "CSafeLoader", yaml.SafeLoader)


def _read_text(filename: str) -> str:
    if filename.endswith(".xz") or filename.endswith(".lzma"):
        opener = lzma.open
    elif filename.endswith(".gz"):
        opener = gzip.open
    elif filename.endswith(".bz2"):
        opener = bz2.open
    else:
        opener = open
    with opener(filename, "rt") as f:
        return f.read()


def _cell_to_structure(cell: dict) -> PolymlpStructure:
    """phonopy cell -> PolymlpStructure (src/pypolymlp/utils/phonopy_utils.py:29-45):
    axis = lattice^T, positions = scaled positions^T, types by first appearance."""
    lattice = np.array(cell["lattice"], dtype=float)
    symbols = [p["symbol"] for p in cell["points"]]
    frac = np.array([p["coordinates"] for p in cell["points"]], dtype=float)
    uniq = sorted(set(symbols), key=symbols.index)
    n_atoms = [symbols.count(e) for e in uniq]
    types = [uniq.index(s) for s in symbols]
    return PolymlpStructure(axis=lattice.T.copy(), positions=frac.T.copy(), n_atoms=n_atoms,
                            elements=symbols, types=types)


class Phono3pyYaml:
    """Perfect supercell, displacements, forces and energies of a phono3py.yaml
    (``Phono3pyYaml``, interface_phono3py.py:75-161). Displacements and forces
    have shape (n_snapshots, 3, n_atoms)."""

    def __init__(self, filename: str, use_phonon_dataset: bool = False):
        doc = yaml.load(_read_text(filename), Loader=_Loader)
        key = "phonon_dataset" if use_phonon_dataset else "dataset"
        cell_key = "phonon_supercell" if use_phonon_dataset and "phonon_supercell" in doc \
            else "supercell"
        if key not in doc or cell_key not in doc:
            raise RuntimeError(f"{filename}: no '{key}' / '{cell_key}' section")
        ds = doc[key]
        if "displacements" not in ds or "forces" not in ds:
            raise RuntimeError(f"{filename}: only datasets with per-atom displacements and "
                               "forces of every snapshot are supported")
        self._supercell = _cell_to_structure(doc[cell_key])
        n = self._supercell.n_total
        disps = np.array(ds["displacements"], dtype=float)
        forces = np.array(ds["forces"], dtype=float)
        if disps.ndim != 3 or disps.shape[1:] != (n, 3) or forces.shape != disps.shape:
            raise RuntimeError(f"{filename}: displacements / forces do not match the supercell")
        self._displacements = disps.transpose((0, 2, 1)).copy()
        self._forces = forces.transpose((0, 2, 1)).copy()
        energies = ds.get("supercell_energies", doc.get("supercell_energies"))
        if energies is None:
            raise RuntimeError("Energy data not found in phono3py.yaml.")
        self._energies = np.array(energies, dtype=float)
        if len(self._energies) != len(disps):
            raise RuntimeError(f"{filename}: {len(self._energies)} energies for "
                               f"{len(disps)} snapshots")
        self._supercells = structures_from_displacements(self._displacements, self._supercell)

    @property
    def supercell(self) -> PolymlpStructure:
        return self._supercell

    @property
    def displacements(self) -> np.ndarray:
        return self._displacements

    @property
    def forces(self) -> np.ndarray:
        return self._forces

    @property
    def energies(self) -> np.ndarray:
        return self._energies

    @property
    def phonon_dataset(self):
        return (self._displacements, self._forces)

    @property
    def supercells(self) -> List[PolymlpStructure]:
        return self._supercells


def structures_from_displacements(disps: np.ndarray, structure: PolymlpStructure):
    """Displaced copies of ``structure`` (displacements.py:10-67): fractional
    positions + inv(axis) @ d for each (3, N) Cartesian displacement d."""
    if disps.shape[1] != 3:
        raise ValueError("displacements must have shape (n, 3, n_atoms)")
    axis_inv = np.linalg.inv(structure.axis)
    out = []
    for i, d in enumerate(disps):
        out.append(PolymlpStructure(axis=structure.axis.copy(),
                                    positions=structure.positions + axis_inv @ d,
                                    n_atoms=list(structure.n_atoms),
                                    elements=list(structure.elements),
                                    types=list(structure.types), name=f"snapshot-{i + 1}"))
    return out


def permute_atoms(st: PolymlpStructure, elements: Sequence[str], force: np.ndarray):
    """Reorder atoms by the element order of the model (types 0, 1, ... in the
    order of ``elements``, stable within an element), as
    ``src/pypolymlp/core/dataset_utils.py:224-229`` does before regression."""
    order, n_atoms, types = [], [], []
    for t, e in enumerate(elements):
        ids = [i for i, s in enumerate(st.elements) if s == e]
        order.extend(ids)
        n_atoms.append(len(ids))
        types.extend([t] * len(ids))
    if len(order) != len(st.elements):
        missing = sorted(set(st.elements) - set(elements))
        raise ValueError(f"elements {missing} of the structure are not in the model")
    st1 = PolymlpStructure(axis=st.axis, positions=st.positions[:, order], n_atoms=n_atoms,
                           elements=[st.elements[i] for i in order], types=types, name=st.name)
    return st1, force[:, order]


def parse_phono3py_yaml(filename: str, elements: Optional[Sequence[str]] = None,
                        use_phonon_dataset: bool = False, return_displacements: bool = False):
    """phono3py.yaml -> Dataset of displaced supercells with energies and forces
    and no stresses (parse_phono3py_yaml, interface_phono3py.py:15-48)."""
    ph3 = Phono3pyYaml(filename, use_phonon_dataset=use_phonon_dataset)
    if elements is None:
        elements = sorted(set(ph3.supercell.elements), key=ph3.supercell.elements.index)
    if not np.all(np.isin(elements, ph3.supercell.elements)):
        raise ValueError("Elements in the input file are not found in phono3py.yaml.")
    sts, fs = [], []
    for st, f in zip(ph3.supercells, ph3.forces):
        st1, f1 = permute_atoms(st, elements, f)
        sts.append(st1)
        fs.append(f1)
    n = len(sts)
    ds = Dataset(structures=sts, energies=ph3.energies.copy(), forces=fs,
                 stresses=np.zeros((n, 6)), name="Data_from_files", exist_stress=False)
    if return_displacements:
        return ds, ph3.displacements
    return ds
