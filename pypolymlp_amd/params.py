"""Model parameters, the params dict, and the polymlp.yaml model format.

The params dict produced here has the schema of
``PolymlpParamsSingle.as_dict()`` (``src/pypolymlp/core/data_format.py:268-272``)
restricted to the keys the hot path reads
(``src/pypolymlp/cxx/src/compute/py_params.cpp:11-60``):

    n_type, print_memory,
    model.{cutoff, pair_type, feature_type, model_type, max_p, max_l,
           gtinv.{order, max_l, version, lm_seq, l_comb, lm_coeffs},
           pair_conditional, pair_params, pair_params_conditional[(j, i)], j <= i}

so that dicts made by pypolymlp itself are accepted unchanged.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import yaml

from .gtinv import get_gtinv_attrs

MASS = {"H": 1.008, "Al": 26.98, "Mg": 24.31, "O": 16.0, "Fe": 55.85, "Cu": 63.55,
        "Ag": 107.9, "Au": 197.0, "Ti": 47.87, "Si": 28.09, "A": 1.0, "B": 1.0, "C": 12.01}


def set_gaussian_params(params1=(1.0, 1.0, 1), params2=(0.0, 5.0, 7)):
    """Gaussian (beta, mu) pairs plus the constant function (params_utils.py:72-94)."""
    g1 = np.linspace(float(params1[0]), float(params1[1]), int(params1[2]))
    g2 = np.linspace(float(params2[0]), float(params2[1]), int(params2[2]))
    pair_params = [list(map(float, p)) for p in itertools.product(g1, g2)]
    pair_params.append([0.0, 0.0])
    return pair_params


def all_pair_conditional(n_type: int, n_params: int):
    """Every radial function active for every type pair (params_utils.py:97-138)."""
    return {
        (i, j): list(range(n_params))
        for i, j in itertools.combinations_with_replacement(range(n_type), 2)
    }


@dataclass
class ModelSpec:
    """Hyperparameters of one polymlp model (PolymlpModelParams subset)."""

    n_type: int
    cutoff: float
    feature_type: str = "gtinv"
    model_type: int = 4
    max_p: int = 2
    gtinv_order: int = 2
    gtinv_maxl: Sequence[int] = (4,)
    gaussian_params1: Sequence[float] = (1.0, 1.0, 1)
    gaussian_params2: Sequence[float] = (0.0, 5.0, 7)
    pair_params: Optional[list] = None
    pair_params_conditional: Optional[dict] = None
    elements: Optional[Sequence[str]] = None
    gtinv_version: int = 1
    extra: dict = field(default_factory=dict)
    # hybrid sub-models: the atom types (of the full system) this model acts
    # on, or every atom when type_full (PolymlpParamsSingle.type_full / type_indices)
    type_full: bool = True
    type_indices: Optional[Sequence[int]] = None

    def __post_init__(self):
        if self.pair_params is None:
            self.pair_params = set_gaussian_params(self.gaussian_params1, self.gaussian_params2)
        if self.pair_params_conditional is None:
            self.pair_params_conditional = all_pair_conditional(self.n_type, len(self.pair_params))
        if self.elements is None:
            self.elements = [chr(ord("A") + i) for i in range(self.n_type)]
        if self.feature_type == "pair":
            self.gtinv_order = 0
            self.gtinv_maxl = ()
        if self.type_indices is None:
            self.type_indices = list(range(self.n_type))

    @property
    def max_l(self) -> int:
        return int(max(self.gtinv_maxl)) if self.feature_type == "gtinv" and self.gtinv_maxl else 0

    def as_dict(self) -> dict:
        """Params dict with the schema of PolymlpParamsSingle.as_dict()."""
        if self.feature_type == "gtinv":
            l_comb, lm_seq, lm_coeffs = get_gtinv_attrs(
                self.gtinv_order, self.gtinv_maxl, self.gtinv_version
            )
        else:
            l_comb, lm_seq, lm_coeffs = [], [], []
        gtinv = {
            "order": self.gtinv_order,
            "max_l": list(self.gtinv_maxl),
            "n_type": self.n_type,
            "version": self.gtinv_version,
            "lm_seq": lm_seq,
            "l_comb": l_comb,
            "lm_coeffs": lm_coeffs,
        }
        model = {
            "cutoff": float(self.cutoff),
            "model_type": int(self.model_type),
            "max_p": int(self.max_p),
            "max_l": self.max_l,
            "feature_type": self.feature_type,
            "gtinv": gtinv,
            "pair_type": "gaussian",
            "pair_conditional": True,
            "pair_params": [list(map(float, p)) for p in self.pair_params],
            "pair_params_conditional": {
                tuple(k): list(v) for k, v in self.pair_params_conditional.items()
            },
        }
        return {
            "n_type": self.n_type,
            "elements": list(self.elements),
            "print_memory": False,
            "include_force": True,
            "include_stress": True,
            "type_full": bool(self.type_full),
            "type_indices": [int(t) for t in self.type_indices],
            "model": model,
        }


def save_mlp_yaml(spec: ModelSpec, coeffs, scales, filename: str):
    """Write polymlp.yaml in the reference layout (io_polymlp_yaml.py:17-86).

    The stored coefficients are ``coeffs / scales`` printed with ``%.15e``.
    """
    coeffs = np.asarray(coeffs, dtype=float)
    scales = np.asarray(scales, dtype=float)
    lines = []
    lines.append("elements:      [" + ", ".join(spec.elements) + "]")
    lines.append("enable_spins:  " + str([0 for _ in spec.elements]))
    lines.append(f"cutoff:        {float(spec.cutoff)}")
    lines.append("pair_type:     gaussian")
    lines.append(f"feature_type:  {spec.feature_type}")
    lines.append(f"model_type:    {spec.model_type}")
    lines.append(f"max_p:         {spec.max_p}")
    lines.append(f"max_l:         {spec.max_l}")
    lines.append("")
    if spec.feature_type == "gtinv":
        lines.append(f"gtinv_order:   {spec.gtinv_order}")
        lines.append(f"gtinv_maxl:    {list(map(int, spec.gtinv_maxl))}")
        lines.append(f"gtinv_sym:     {[0 for _ in spec.gtinv_maxl]}")
        lines.append(f"gtinv_version: {spec.gtinv_version}")
        lines.append("")
    lines.append("electrostatic: 0")
    lines.append("mass:          " + str([MASS.get(e, 1.0) for e in spec.elements]))
    lines.append("")
    lines.append(f"n_pair_params: {len(spec.pair_params)}")
    lines.append("pair_params:")
    for p in spec.pair_params:
        lines.append("- " + str([float(p[0]), float(p[1])]))
    lines.append("")
    lines.append(f"n_type_pairs: {len(spec.pair_params_conditional)}")
    lines.append("type_pairs:")
    for k, v in spec.pair_params_conditional.items():
        lines.append("- atom_type_pair:      " + str(list(k)))
        lines.append("  pair_params_indices: " + str(list(v)))
    lines.append("")
    lines.append(f"type_full:    {int(bool(spec.type_full))}")
    lines.append("type_indices: " + str([int(t) for t in spec.type_indices]))
    lines.append("")
    c = coeffs / scales
    lines.append(f"n_coeffs: {len(c)}")
    lines.append("coeffs: [" + ", ".join(f"{v:.15e}" for v in c) + "]")
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")


def load_mlp_yaml(filename: str):
    """Read polymlp.yaml (io_polymlp_yaml.py:89-154). Returns (ModelSpec, coeffs)."""
    yml = yaml.safe_load(open(filename))
    elements = list(yml["elements"])
    n_type = len(elements)
    # spin-resolved models split each spin-enabled element into two atom types
    # (PolymlpParams._set_params_spins, data_format.py:291-320) and need spin
    # labels on the atoms; this build does not implement them, so refuse them
    # rather than evaluate them with the element types
    spins = [int(v) for v in yml.get("enable_spins", [0] * n_type) or []]
    if any(spins):
        raise ValueError(f"{filename}: spin-resolved models (enable_spins {spins}) are not "
                         "supported")
    cond = {tuple(tp["atom_type_pair"]): list(tp["pair_params_indices"]) for tp in yml["type_pairs"]}
    ft = yml["feature_type"]
    spec = ModelSpec(
        n_type=n_type,
        cutoff=float(yml["cutoff"]),
        feature_type=ft,
        model_type=int(yml["model_type"]),
        max_p=int(yml["max_p"]),
        gtinv_order=int(yml["gtinv_order"]) if ft == "gtinv" else 0,
        gtinv_maxl=tuple(yml["gtinv_maxl"]) if ft == "gtinv" else (),
        pair_params=[list(map(float, p)) for p in yml["pair_params"]],
        pair_params_conditional=cond,
        elements=elements,
        gtinv_version=int(yml.get("gtinv_version", 1)) if ft == "gtinv" else 1,
        type_full=bool(yml.get("type_full", 1)),
        type_indices=list(yml.get("type_indices", range(n_type))),
    )
    return spec, np.array(yml["coeffs"], dtype=float)


# --- polymlp.in ------------------------------------------------------------
@dataclass
class PolymlpInput:
    """Settings of a polymlp.in file that the fit path reads
    (ParamsParser, src/pypolymlp/core/parser_polymlp_params.py:20-180)."""

    spec: ModelSpec
    alphas: np.ndarray
    atomic_energy: tuple
    include_force: bool = True
    include_stress: bool = True
    dataset_type: str = "vasp"
    data: list = field(default_factory=list)  # [(tag, [strings])], tag in data/train_data/test_data


def _tobool(s: str) -> bool:
    if s.lower() in ("true", "t", "yes", "y", "1", ".true."):
        return True
    if s.lower() in ("false", "f", "no", "n", "0", ".false."):
        return False
    raise ValueError(f"not a boolean: {s}")


def parse_polymlp_in(filename: str) -> PolymlpInput:
    """Read a polymlp.in (one 'tag values...' per line, '#' comments) with the
    reference's defaults: n_type 1, cutoff 6, model_type 1, max_p 1,
    feature_type gtinv, gtinv_order 3, gtinv_maxl padded with 2 to order - 1
    entries, gaussian_params1 (1, 1, 1), gaussian_params2 (0, cutoff - 1, 7),
    reg_alpha_params (-3, 1, 5) giving alphas 10^linspace, atomic energies 0."""
    tags, data = {}, []
    for line in open(filename):
        line = line.split("#")[0].split()
        if not line:
            continue
        if line[0] in ("data", "train_data", "test_data", "data_md"):
            data.append((line[0], line[1:]))
        else:
            tags[line[0]] = line[1:]
    get = lambda k, d: tags.get(k, d)  # noqa: E731
    n_type = int(get("n_type", ["1"])[0])
    elements = get("elements", None)
    cutoff = float(get("cutoff", ["6.0"])[0])
    ftype = get("feature_type", ["gtinv"])[0]
    order = int(get("gtinv_order", ["3"])[0])
    maxl = [int(v) for v in get("gtinv_maxl", [])]
    maxl += [2] * max(0, order - 1 - len(maxl))
    p1 = [float(v) for v in get("gaussian_params1", ["1.0", "1.0", "1"])]
    p2 = [float(v) for v in get("gaussian_params2", ["0.0", str(cutoff - 1.0), "7"])]
    a = [float(v) for v in get("reg_alpha_params", ["-3", "1", "5"])]
    atom_e = tuple(float(v) for v in get("atomic_energy", ["0.0"] * n_type))
    include_force = _tobool(get("include_force", ["True"])[0])
    include_stress = _tobool(get("include_stress", ["True"])[0]) if include_force else False
    dataset_type = get("dataset_type", ["vasp"])[0]
    # dataset types without stress data fit without stress rows
    # (PolymlpParams._set_tags, src/pypolymlp/core/params.py:369-383)
    if dataset_type in ("phono3py", "openmx"):
        include_stress = False
    elif dataset_type == "electron":
        include_force = include_stress = False
    spec = ModelSpec(n_type=n_type, cutoff=cutoff, feature_type=ftype,
                     model_type=int(get("model_type", ["1"])[0]),
                     max_p=int(get("max_p", ["1"])[0]), gtinv_order=order,
                     gtinv_maxl=tuple(maxl[:order - 1]),
                     gaussian_params1=(p1[0], p1[1], int(p1[2])),
                     gaussian_params2=(p2[0], p2[1], int(p2[2])), elements=elements,
                     gtinv_version=int(get("gtinv_version", ["1"])[0]))
    alphas = np.array([10 ** x for x in np.linspace(a[0], a[1], int(a[2]))])
    return PolymlpInput(spec=spec, alphas=alphas, atomic_energy=atom_e,
                        include_force=include_force, include_stress=include_stress,
                        dataset_type=dataset_type, data=data)


# --- legacy polymlp.lammps -------------------------------------------------
def _legacy_var(line: str, dtype=int, return_list: bool = False):
    vals = line.split("#")[0].split()
    if return_list:
        return [dtype(v) for v in vals]
    return dtype(vals[0])


def load_mlp_lammps(filename: str):
    """Read the legacy polymlp.lammps format (load_mlp_lammps,
    src/pypolymlp/core/io_polymlp_legacy.py:23-161): one value or list per
    line with '#' comments, in the order elements, cutoff, pair_type,
    feature_type, model_type, max_p, max_l, [gtinv_order, gtinv_maxl,
    gtinv_sym], n_coeffs, coeffs, scales, n_params, (beta, mu) per line, mass,
    electrostatic, [gtinv_version], [n_type_pairs and (type pair, indices)
    line pairs], [type_full, type_indices]. Returns (ModelSpec, coeffs / scales);
    without a type-pair section every radial function serves every pair."""
    lines = open(filename).read().splitlines()
    i = 0

    def nxt(dtype=int, return_list=False):
        nonlocal i
        v = _legacy_var(lines[i], dtype, return_list)
        i += 1
        return v

    elements = nxt(str, True)
    n_type = len(elements)
    cutoff = nxt(float)
    nxt(str)  # pair_type (gaussian)
    ftype = nxt(str)
    model_type = nxt()
    max_p = nxt()
    nxt()  # max_l (derived from gtinv_maxl)
    order, maxl = 0, ()
    if ftype == "gtinv":
        order = nxt()
        maxl = tuple(nxt(int, True))
        nxt(_tobool, True)  # gtinv_sym
    nxt()  # n_coeffs
    coeffs = np.array(nxt(float, True))
    scales = np.array(nxt(float, True))
    n_pp = nxt()
    pair_params = [nxt(float, True) for _ in range(n_pp)]
    nxt(float, True)  # mass
    nxt(_tobool)  # electrostatic
    version = 1
    if ftype == "gtinv" and i < len(lines) and "gtinv_version" in lines[i]:
        version = nxt()
    cond = None
    if i < len(lines) and "n_type_pairs" in lines[i]:
        cond = {}
        for _ in range(nxt()):
            tp = tuple(nxt(int, True)[0:2])
            cond[tp] = nxt(int, True)
    type_full, type_indices = True, list(range(n_type))
    if i < len(lines) and "type_full" in lines[i]:
        type_full = nxt(_tobool)
        type_indices = nxt(int, True)
    spec = ModelSpec(n_type=n_type, cutoff=cutoff, feature_type=ftype, model_type=model_type,
                     max_p=max_p, gtinv_order=order, gtinv_maxl=maxl, pair_params=pair_params,
                     pair_params_conditional=cond, elements=elements, gtinv_version=version,
                     type_full=type_full, type_indices=type_indices)
    return spec, coeffs / scales


def load_mlp(filename: str):
    """polymlp.yaml or legacy polymlp.lammps, told apart by the first line
    (the yaml format starts with an 'elements:' key)."""
    with open(filename) as f:
        first = f.readline()
    if first.lstrip().startswith("elements:"):
        return load_mlp_yaml(filename)
    return load_mlp_lammps(filename)
