"""Hybrid models: several polymlp models whose energies add up.

Host-side mirror of the reference's hybrid path:
  PyHybridModel (X)       src/pypolymlp/cxx/src/compute/py_hybrid_model.cpp:11-143
  FeaturesHybrid          src/pypolymlp/mlp_dev/core/features.py:172-243
  PropertiesHybrid (eval) src/pypolymlp/calculator/properties_hybrid.py:13-78,
                          properties_single.py:43-150, utils/properties_utils.py:9-44

X of a hybrid model is the column concatenation of the sub-models' X. A
sub-model with ``type_full`` acts on every atom; otherwise only on the atoms
whose type is in ``type_indices``, renumbered by their position there, and
its force rows of the other atoms are zero. Every sub-model's X and
evaluation run on the GPU through the same batch kernels as a single model;
only the selection of active atoms and the scatter of their force rows are
host work.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .datasets import Dataset
from .mlp_dev import Features, get_min_energy, round_scales, row_weights, DataXY
from .params import ModelSpec
from .structure import PolymlpStructure


@dataclass
class SubModel:
    spec: ModelSpec
    model: capi.Model

    @classmethod
    def from_spec(cls, spec: ModelSpec, coeffs: Optional[np.ndarray] = None) -> "SubModel":
        m = capi.Model(spec.as_dict())
        if coeffs is not None:
            m.set_coeffs(np.asarray(coeffs, dtype=float))
        return cls(spec=spec, model=m)


def active_atoms(st: PolymlpStructure, spec: ModelSpec, by_elements: bool = False):
    """(atom indices, renumbered types) of the atoms the sub-model acts on.

    By type (training, PyHybridModel::find_active_atoms,
    py_hybrid_model.cpp:107-143): atoms whose type is in type_indices, typed
    by position there. By element (evaluation, find_active_atoms and
    update_types, properties_utils.py:9-44, compute_features.py:15-40): atoms
    whose element is one of the model's, typed by the model's element order."""
    types = np.asarray(st.types)
    if by_elements and st.elements is not None and len(st.elements) == len(types):
        ele = list(spec.elements)
        if spec.type_full and not set(st.elements) <= set(ele):
            return np.arange(len(types)), types
        sel = np.array([i for i, e in enumerate(st.elements) if e in ele], dtype=int)
        return sel, np.array([ele.index(st.elements[i]) for i in sel], dtype=int)
    if spec.type_full:
        return np.arange(len(types)), types
    idx = list(spec.type_indices)
    sel = np.array([i for i, t in enumerate(types) if t in idx], dtype=int)
    return sel, np.array([idx.index(int(types[i])) for i in sel], dtype=int)


def _sub_batch(sub: SubModel, structures: Sequence[PolymlpStructure], by_elements=False):
    """Batch of the active parts of the structures that have active atoms."""
    keep, atoms, axes, pos, types = [], [], [], [], []
    for k, st in enumerate(structures):
        sel, t = active_atoms(st, sub.spec, by_elements)
        if len(sel) == 0:
            continue
        keep.append(k)
        atoms.append(sel)
        axes.append(st.axis)
        pos.append(st.positions_cartesian[:, sel])
        types.append(t)
    batch = capi.Batch(sub.model, axes, pos, types) if keep else None
    return keep, atoms, batch


def compute_features_hybrid(subs: Sequence[SubModel], datasets: Sequence[Dataset],
                            precision=capi.PLM_FP64) -> Features:
    """X (rows x sum F_n) in the layout of compute_features (PyHybridModel,
    py_hybrid_model.cpp:45-103): energy rows, then 6 stress rows and 3N force
    rows per structure of the datasets with forces; sub-model n fills columns
    [cum_n, cum_{n+1}), its force rows scattered to the active atoms."""
    sts = [s for d in datasets for s in d.structures]
    force_st = np.concatenate([[d.include_force] * len(d.structures) for d in datasets])
    n_st = len(sts)
    n_at = np.array([s.n_total for s in sts])
    ns_tot = 6 * int(force_st.sum())
    s_begin = n_st + 6 * (np.cumsum(force_st) - force_st)
    f_cnt = np.where(force_st, 3 * n_at, 0)
    f_begin = n_st + ns_tot + np.cumsum(f_cnt) - f_cnt
    n_rows = n_st + ns_tot + int(f_cnt.sum())
    cum = np.cumsum([s.model.n_features for s in subs])
    X = np.zeros((n_rows, int(cum[-1])))
    for n, sub in enumerate(subs):
        c0 = int(cum[n - 1]) if n else 0
        keep, atoms, batch = _sub_batch(sub, sts)
        if batch is None:
            continue
        Xs = capi.features_x(sub.model, batch, precision)
        Fn = sub.model.n_features
        m = len(keep)
        na = np.array([len(a) for a in atoms])
        foff = m * 7 + np.concatenate([[0], np.cumsum(3 * na)])
        for q, (k, sel) in enumerate(zip(keep, atoms)):
            X[k, c0:c0 + Fn] = Xs[q]
            if force_st[k]:
                X[s_begin[k]:s_begin[k] + 6, c0:c0 + Fn] = Xs[m + 6 * q:m + 6 * q + 6]
                rows = f_begin[k] + (3 * sel[:, None] + np.arange(3)[None, :]).reshape(-1)
                X[rows, c0:c0 + Fn] = Xs[foff[q]:foff[q + 1]]
    first, i0, ns, nf = [], 0, n_st, n_st + ns_tot
    for d in datasets:
        n = len(d.structures)
        if d.include_force:
            first.append((i0, nf, ns))
            ns += 6 * n
            nf += 3 * int(n_at[i0:i0 + n].sum())
        else:
            first.append((i0, -1, -1))
        i0 += n
    feat = Features(x=X, first_indices=first, n_data=(n_st, int(f_cnt.sum()), ns_tot))
    feat.cumulative_n_features = [int(c) for c in cum]
    return feat


def calc_xy_hybrid(subs: Sequence[SubModel], datasets: Sequence[Dataset], scales=None,
                   min_energy=None, weight_stress: float = 0.1, scale_threshold: float = 1e-10,
                   precision=capi.PLM_FP64) -> DataXY:
    """calc_xy (data_standard.py:16-77) over the hybrid X."""
    feat = compute_features_hybrid(subs, datasets, precision)
    x = feat.x
    ne, nf, _ = feat.n_data
    if scales is None:
        scales = np.std(x[:ne], axis=0)
    scales, zero = round_scales(np.array(scales, dtype=float), include_force=nf > 0,
                                threshold=scale_threshold)
    x[:, zero] = 0.0
    x /= scales
    if min_energy is None:
        min_energy = get_min_energy(datasets)
    y = np.zeros(x.shape[0])
    w = np.ones(x.shape[0])
    for d, (eb, fb, sb) in zip(datasets, feat.first_indices):
        wd, yd = row_weights(d, min_energy, weight_stress)
        n = len(d.energies)
        rows = [np.arange(eb, eb + n)]
        if d.include_force:
            rows += [np.arange(sb, sb + 6 * n), np.arange(fb, fb + 3 * int(d.total_n_atoms.sum()))]
        else:
            wd, yd = wd[:n], yd[:n]
        rows = np.concatenate(rows)
        w[rows] = wd
        y[rows] = yd
        x[rows] *= wd[:, None]
    out = DataXY(x=x, y=y, weights=w, scales=scales, min_energy=float(min_energy),
                 first_indices=feat.first_indices, n_structures=ne)
    out.cumulative_n_features = feat.cumulative_n_features
    return out


def split_coeffs(coeffs, subs: Sequence[SubModel]) -> List[np.ndarray]:
    """Coefficients of a hybrid fit, one array per sub-model."""
    cum = np.cumsum([0] + [s.model.n_features for s in subs])
    return [np.asarray(coeffs[cum[i]:cum[i + 1]]) for i in range(len(subs))]


def eval_hybrid(subs: Sequence[SubModel], structures: Sequence[PolymlpStructure],
                precision=capi.PLM_FP64):
    """(energies, forces [(N, 3)], stresses (n, 6)): the sum of the sub-models'
    evaluations, each over its active atoms (selected by element when the
    structures carry element names), forces scattered back
    (PropertiesHybrid.eval_multiple with PropertiesSingle.eval_multiple).
    Each sub-model's coefficients must be set on its Model."""
    n = len(structures)
    E = np.zeros(n)
    F = [np.zeros((s.n_total, 3)) for s in structures]
    S = np.zeros((n, 6))
    for sub in subs:
        keep, atoms, batch = _sub_batch(sub, structures, by_elements=True)
        if batch is None:
            continue
        e, f, s = capi.eval_batch(sub.model, batch, precision)
        for q, (k, sel) in enumerate(zip(keep, atoms)):
            E[k] += e[q]
            F[k][sel] += f[q]
            S[k] += s[q]
    return E, F, S
