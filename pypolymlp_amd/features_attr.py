"""Attributes of the columns of X (get_features_attr, get_num_features,
src/pypolymlp/mlp_dev/core/features_attr.py:12-49) from the library's host
tables (plm_model_feature_attrs, the counterpart of FeaturesAttr,
src/pypolymlp/cxx/src/compute/py_features_attr.cpp:11-44)."""

from __future__ import annotations

from collections import defaultdict
from typing import Sequence

from . import capi
from .params import ModelSpec


def get_features_attr(spec: ModelSpec):
    """(features_attr, polynomial_attr, atomtype_pair_dict): per linear feature
    (radial id, [gtinv id,] type-pair combination); per polynomial column its
    linear features; for each type pair the (type1, type2) that map to it."""
    radial, gtinv, tcomb, poly, type_pairs = capi.feature_attrs(capi.Model(spec.as_dict()))
    pairs = defaultdict(list)
    for t1, tps in enumerate(type_pairs):
        for t2, tp in enumerate(tps):
            pairs[tp].append([t1, t2])
    if spec.feature_type == "pair":
        feats = list(zip(radial, tcomb))
    else:
        feats = list(zip(radial, gtinv, tcomb))
    return feats, poly, pairs


def get_num_features(specs: Sequence[ModelSpec]) -> int:
    """Columns of X of a (hybrid) model: linear features plus polynomial columns."""
    if isinstance(specs, ModelSpec):
        specs = [specs]
    n = 0
    for s in specs:
        feats, poly, _ = get_features_attr(s)
        n += len(feats) + len(poly)
    return n
