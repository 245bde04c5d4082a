"""Coupling-coefficient tables of the O(3) polynomial invariants.

``Readgtinv`` mirrors the reference class of the same name
(``src/pypolymlp/cxx/src/polymlp/polymlp_read_gtinv.cpp:12-59``): for
invariant orders 1..order it keeps the l-combinations whose last (largest) l is
at most ``gtinv_maxl[order-2]`` and converts each (l, m) of a term into the
flat index ``l*l + l + m``. ``get_gtinv_attrs`` mirrors
``src/pypolymlp/cxx/wrapper/api_gtinv_list.py:6-12`` (returns
``l_comb, lm_seq, lm_coeffs``).

The table values are read from ``data/gtinv_v1.npz`` (converted from the
reference's version-1 binary tables by ``data/convert_gtinv_tables.py``).
Only version 1 is shipped: the reference's version-2 tables for orders 3-6
are not present in its tree.
"""

from __future__ import annotations

import os
from functools import lru_cache

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "gtinv_v1.npz")


@lru_cache(maxsize=1)
def _tables():
    with np.load(_DATA) as d:
        out = {}
        for order in range(1, 7):
            l_all = d[f"l{order}"].astype(np.int64)
            nt = d[f"nt{order}"].astype(np.int64)
            m_all = d[f"m{order}"].astype(np.int64)
            c_all = d[f"c{order}"]
            off = np.concatenate([[0], np.cumsum(nt)])
            out[order] = (l_all, off, m_all, c_all)
        return out


class Readgtinv:
    """Screened invariant tables (polymlp_read_gtinv.cpp:23-59)."""

    def __init__(self, gtinv_order: int, gtinv_maxl, version: int = 1):
        if version != 1:
            raise RuntimeError("Only version-1 invariant tables are available.")
        if gtinv_order < 1 or gtinv_order > 6:
            raise ValueError("Invalid order")
        maxl = list(gtinv_maxl)
        if len(maxl) < gtinv_order - 1:
            raise ValueError("gtinv_maxl must have gtinv_order - 1 entries")
        tables = _tables()
        self._l, self._lm, self._c = [], [], []
        for order in range(1, gtinv_order + 1):
            l_all, off, m_all, c_all = tables[order]
            lmax = 0 if order == 1 else int(maxl[order - 2])
            for i in range(l_all.shape[0]):
                lcomb = l_all[i]
                if lmax < lcomb[-1]:
                    continue
                m = m_all[off[i] : off[i + 1]]
                lm = lcomb[None, :] * lcomb[None, :] + lcomb[None, :] + m
                self._l.append([int(v) for v in lcomb])
                self._lm.append(lm.astype(np.int64).tolist())
                self._c.append(c_all[off[i] : off[i + 1]].tolist())

    def get_lm_seq(self):
        return self._lm

    def get_l_comb(self):
        return self._l

    def get_lm_coeffs(self):
        return self._c


def get_gtinv_attrs(order: int, lmax, version: int = 1):
    """Return (l_comb, lm_seq, lm_coeffs) as in api_gtinv_list.py:6-12."""
    rgt = Readgtinv(order, lmax, version)
    return rgt.get_l_comb(), rgt.get_lm_seq(), rgt.get_lm_coeffs()
