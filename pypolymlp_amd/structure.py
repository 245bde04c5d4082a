"""Crystal structures: PolymlpStructure subset and a POSCAR reader.

Conventions follow ``src/pypolymlp/core/data_format.py`` and
``src/pypolymlp/core/interface_vasp.py:380-420``: ``axis`` holds the lattice
vectors as *columns*, ``positions`` are fractional coordinates (3 x N),
``types`` are integer atom types and the Cartesian positions handed to the
hot path are ``axis @ positions`` (``src/pypolymlp/mlp_dev/core/features.py:16``),
with no rewrapping.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np


@dataclass
class PolymlpStructure:
    axis: np.ndarray  # (3, 3), lattice vectors as columns
    positions: np.ndarray  # (3, N) fractional
    n_atoms: List[int]
    elements: List[str]
    types: List[int]
    name: Optional[str] = None

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.axis)))

    @property
    def positions_cartesian(self) -> np.ndarray:
        return self.axis @ self.positions

    @property
    def n_total(self) -> int:
        return int(sum(self.n_atoms))


def read_poscar(filename: str) -> PolymlpStructure:
    """Read a VASP POSCAR (direct coordinates)."""
    lines = [ln for ln in open(filename).read().splitlines()]
    scale = float(lines[1].split()[0])
    axis = np.array([[float(x) for x in lines[i].split()[:3]] for i in (2, 3, 4)]).T * scale
    names = lines[5].split()
    counts = [int(x) for x in lines[6].split()]
    start = 8 if lines[7].strip()[0] in "DdCc" else 9
    cart = lines[start - 1].strip()[0] in "Cc"
    n = sum(counts)
    pos = np.array([[float(x) for x in lines[start + i].split()[:3]] for i in range(n)]).T
    if cart:
        pos = np.linalg.solve(axis, pos * scale)
    elements, types = [], []
    for t, (name, c) in enumerate(zip(names, counts)):
        elements.extend([name] * c)
        types.extend([t] * c)
    return PolymlpStructure(axis=axis, positions=pos, n_atoms=counts, elements=elements, types=types)


def volume_gpa_factor(structure: PolymlpStructure) -> float:
    """Factor converting eV/cell stress to GPa (properties_utils.py:45-60)."""
    ev_to_gpa = 160.21766208
    return ev_to_gpa / structure.volume
