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


def supercell(st: PolymlpStructure, matrix) -> PolymlpStructure:
    """Supercell for an integer matrix H with lattice vectors a'_i = sum_j H_ij a_j:
    the images of every atom inside the new cell, fractional coordinates
    wrapped to [0, 1), atoms grouped by type (for non-diagonal H the reference
    builds it with phonopy, src/pypolymlp/utils/structure_utils.py:83-100)."""
    H = np.array(matrix, dtype=int)
    if H.ndim == 1:
        H = np.diag(H)
    det = int(round(abs(np.linalg.det(H))))
    if det == 0:
        raise ValueError("singular supercell matrix")
    axis = st.axis @ H.T
    inv = np.linalg.inv(H.T.astype(float))
    # lattice points of the old cell inside the new one: search the bounding box
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)]) @ H
    lo, hi = corners.min(axis=0), corners.max(axis=0)
    pts = []
    for i in range(lo[0], hi[0] + 1):
        for j in range(lo[1], hi[1] + 1):
            for k in range(lo[2], hi[2] + 1):
                f = inv @ np.array([i, j, k], dtype=float)
                if np.all(f >= -1e-10) and np.all(f < 1 - 1e-10):
                    pts.append((i, j, k))
    if len(pts) != det:
        raise RuntimeError("lattice-point search failed")
    frac, types, elements = [], [], []
    order = np.argsort(st.types, kind="stable")
    for a in order:
        for p in pts:
            f = inv @ (st.positions[:, a] + np.array(p, dtype=float))
            frac.append(f - np.floor(f))
            types.append(int(st.types[a]))
            elements.append(st.elements[a])
    n_atoms = [types.count(t) for t in range(max(types) + 1)]
    return PolymlpStructure(axis=axis, positions=np.array(frac).T, n_atoms=n_atoms,
                            elements=elements, types=types)
