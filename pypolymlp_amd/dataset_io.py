"""Training and test datasets from a polymlp.in
(ParamsParser.set_datasets, src/pypolymlp/core/parser_polymlp_params.py:229-280;
Dataset, src/pypolymlp/core/dataset.py:22-140).

- phono3py: a 'data' entry names one phono3py.yaml; it is parsed, its atomic
  energies subtracted, and then split 90 / 10 by structure
  (Dataset.split_dft, DatasetDFT.split).
- vasp: an entry is 'location [include_force [weight]]', a glob of
  vasprun.xml files (sorted); a 'data' entry is split 90 / 10 by file before
  parsing (Dataset.split_files), a 'data_md' entry by structure after it.
'train_data' / 'test_data' entries are taken whole.
"""

from __future__ import annotations

import glob
import os
from typing import List, Optional, Tuple

from .datasets import Dataset, split_ids_train_test
from .params import PolymlpInput, _tobool


def _finish(ds: Dataset, inp: PolymlpInput, include_force: bool, weight: float) -> Dataset:
    ds.include_force = include_force
    ds.include_stress = inp.include_stress if include_force else False
    ds.weight = weight
    return ds.apply_atomic_energy(inp.atomic_energy)


def _vasp_entry(inp: PolymlpInput, tag: str, strings, prefix: Optional[str], train_ratio: float):
    from .interface_vasp import set_dataset_from_vaspruns
    loc = strings[0] if prefix is None else os.path.join(prefix, strings[0])
    files = sorted(glob.glob(loc))
    if not files:
        raise FileNotFoundError(f"no files match {loc}")
    include_force = _tobool(strings[1]) if len(strings) > 1 else inp.include_force
    include_force = include_force and inp.include_force
    weight = float(strings[2]) if len(strings) > 2 else 1.0
    if tag == "data":
        tr, te = split_ids_train_test(len(files), train_ratio=train_ratio)
        parts = [("Train_" + strings[0], [files[i] for i in tr]),
                 ("Test_" + strings[0], [files[i] for i in te])]
    else:
        parts = [(strings[0], files)]
    out = []
    for name, fl in parts:
        if not fl:
            out.append(None)
            continue
        ds = set_dataset_from_vaspruns(fl, elements=inp.spec.elements, name=name)
        out.append(_finish(ds, inp, include_force, weight))
    return out


def _phono3py_entry(inp: PolymlpInput, strings, prefix: Optional[str]) -> Dataset:
    from .interface_phono3py import parse_phono3py_yaml
    path = strings[0] if prefix is None else os.path.join(prefix, strings[0])
    ds = parse_phono3py_yaml(path, elements=inp.spec.elements)
    return _finish(ds, inp, inp.include_force, 1.0)


def load_datasets(inp: PolymlpInput, prefix: Optional[str] = None,
                  train_ratio: float = 0.9) -> Tuple[List[Dataset], List[Dataset]]:
    train, test = [], []
    for tag, strings in inp.data:
        if inp.dataset_type == "vasp":
            parts = _vasp_entry(inp, tag, strings, prefix, train_ratio)
            if tag == "data":
                if parts[0] is not None:
                    train.append(parts[0])
                if parts[1] is not None:
                    test.append(parts[1])
            elif tag == "data_md":
                tr, te = parts[0].split(train_ratio)
                if tr is not None:
                    train.append(tr)
                if te is not None:
                    test.append(te)
            else:
                (train if tag == "train_data" else test).append(parts[0])
            continue
        if inp.dataset_type != "phono3py":
            raise NotImplementedError(f"dataset_type {inp.dataset_type} is not supported "
                                      "(supported: vasp, phono3py)")
        ds = _phono3py_entry(inp, strings, prefix)
        if tag == "train_data":
            train.append(ds)
        elif tag == "test_data":
            test.append(ds)
        else:
            tr, te = ds.split(train_ratio)
            if tr is not None:
                train.append(tr)
            if te is not None:
                test.append(te)
    return train, test
