"""Reader of VASP vasprun.xml files of single-point calculations.

Host-side mirror of ``src/pypolymlp/core/interface_vasp.py``:
  set_dataset_from_vaspruns / parse_properties_from_vaspruns   :15-75
  check_vasprun_type                                            :93-105
  Vasprun (energy e[1] of the last calculation's <energy>, forces,
  stress in kbar, final structure)                              :108-237
  _is_convergent (|dE| of the last two SC steps < 1e-3)        :516-532

Stresses are converted to eV per cell, sigma * volume / 1602.1766208.
Molecular-dynamics runs (IBRION = 0) are not read.
"""

from __future__ import annotations

import xml.etree.ElementTree as ET
from typing import List, Optional, Sequence

import numpy as np

from .datasets import Dataset
from .interface_phono3py import permute_atoms
from .structure import PolymlpStructure

EV_TO_KBAR = 1602.1766208


def _varray(v) -> np.ndarray:
    return np.array([[float(x) for x in row.text.split()] for row in v])


def _rc_set(rcs):
    return [[c.text.replace(" ", "") for c in rc.findall("c")] for rc in rcs]


def check_vasprun_type(name: str):
    """(is an MD run, root): IBRION = 0 marks molecular dynamics."""
    root = ET.parse(name).getroot()
    tag = root.find(".//*[@name='IBRION']")
    try:
        return int(tag.text) == 0, root
    except (AttributeError, TypeError, ValueError):
        return False, root


class Vasprun:
    """Energy, forces (3, N), stress (3, 3, kbar) and final structure of a
    single-point vasprun.xml."""

    def __init__(self, name: str, root=None):
        self._root = ET.parse(name).getroot() if root is None else root
        self._calc = self._root.find("calculation")
        self._name = str(name)

    @property
    def energy(self) -> float:
        return float(self._calc.find("energy")[1].text)

    @property
    def forces(self) -> np.ndarray:
        return _varray(self._root.find(".//*[@name='forces']")).T

    @property
    def stress(self) -> np.ndarray:
        return _varray(self._root.find(".//*[@name='stress']"))

    @property
    def structure(self) -> PolymlpStructure:
        st = self._root.find(".//*[@name='finalpos']")
        axis = _varray(st.find(".//*[@name='basis']")).T
        positions = _varray(st.find(".//*[@name='positions']")).T
        n_atoms = [int(r[0]) for r in _rc_set(self._root.findall(".//*[@name='atomtypes']/set/rc"))]
        atoms = _rc_set(self._root.findall(".//*[@name='atoms']/set/rc"))
        elements = ["Zr" if r[0] == "r" else r[0] for r in atoms]
        types = [int(r[1]) - 1 for r in atoms]
        return PolymlpStructure(axis=axis, positions=positions, n_atoms=n_atoms,
                                elements=elements, types=types, name=self._name)

    def scsteps(self) -> np.ndarray:
        out = []
        for sc in self._calc.findall("scstep"):
            out.append(float(sc.find("energy").find(".//*[@name='e_0_energy']").text))
        return np.array(out)


def is_convergent(v: Vasprun, tol: float = 1e-3) -> bool:
    """The last two self-consistency energies differ by less than tol (true
    when the file holds fewer than two)."""
    e = v.scsteps()
    if len(e) < 2:
        return True
    return abs(e[-1] - e[-2]) < tol


def parse_properties_from_vaspruns(files: Sequence[str]):
    """(structures, (energies, forces [(3, N)], stresses [(3, 3) eV/cell])) of the
    readable, convergent single-point files, in order."""
    sts, es, fs, ss = [], [], [], []
    for name in files:
        md, root = check_vasprun_type(name)
        if md:
            raise NotImplementedError(f"{name}: molecular-dynamics vasprun files are not read")
        v = Vasprun(name, root)
        if not is_convergent(v):
            continue
        st = v.structure
        sts.append(st)
        es.append(v.energy)
        fs.append(v.forces)
        ss.append(v.stress * st.volume / EV_TO_KBAR)
    return sts, (np.array(es), fs, ss)


def set_dataset_from_vaspruns(files: Sequence[str], elements: Optional[Sequence[str]] = None,
                              name: str = "Data_from_files") -> Dataset:
    """Dataset with stresses as (xx, yy, zz, xy, yz, zx) per structure
    (DatasetDFT._set_properties, dataset_utils.py:200-262), atoms ordered by the
    model's elements."""
    sts, (es, fs, ss) = parse_properties_from_vaspruns(files)
    out_st, out_f = [], []
    for st, f in zip(sts, fs):
        if elements is not None:
            st, f = permute_atoms(st, elements, f)
        out_st.append(st)
        out_f.append(f)
    stress6 = np.array([[s[0][0], s[1][1], s[2][2], s[0][1], s[1][2], s[2][0]] for s in ss])
    return Dataset(structures=out_st, energies=np.asarray(es, dtype=float), forces=out_f,
                   stresses=stress6.reshape(-1, 6), name=name)
