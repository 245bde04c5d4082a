"""Training and test datasets from a polymlp.in
(ParamsParser.set_datasets, src/pypolymlp/core/parser_polymlp_params.py:229-280).

A 'data' entry whose source cannot be split by file (a single phono3py.yaml)
is parsed, its atomic energies subtracted, and then split 90 / 10 by
structure (Dataset.split_dft); 'train_data' / 'test_data' entries are taken
whole. Supported dataset types: phono3py.
"""

from __future__ import annotations

import os
from typing import List, Optional, Tuple

from .datasets import Dataset
from .params import PolymlpInput


def _parse(inp: PolymlpInput, strings, prefix: Optional[str]) -> Dataset:
    if inp.dataset_type != "phono3py":
        raise NotImplementedError(f"dataset_type {inp.dataset_type} is not supported "
                                  "(supported: phono3py)")
    from .interface_phono3py import parse_phono3py_yaml
    path = strings[0] if prefix is None else os.path.join(prefix, strings[0])
    ds = parse_phono3py_yaml(path, elements=inp.spec.elements)
    ds.include_force = inp.include_force
    ds.include_stress = inp.include_stress
    return ds.apply_atomic_energy(inp.atomic_energy)


def load_datasets(inp: PolymlpInput, prefix: Optional[str] = None,
                  train_ratio: float = 0.9) -> Tuple[List[Dataset], List[Dataset]]:
    train, test = [], []
    for tag, strings in inp.data:
        ds = _parse(inp, strings, prefix)
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
