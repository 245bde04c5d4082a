"""Convert the reference's coupling-coefficient tables into gtinv_v1.npz.

The O(3) invariants of pypolymlp are fixed linear combinations of products of
order parameters. Their coefficients are data tables that the reference ships
as little-endian binary files ``polymlp_gtinv_data_v1_order{1..6}.bin``
(format written by ``src/pypolymlp/polyinv/binary/convert_binary.py:1-93`` and
read by ``src/pypolymlp/cxx/src/polymlp/polymlp_gtinv_binary.cpp:11-90``):

    b"DATA", int32 n_blocks,
    int32 type=1, 2-D int32   l-combinations     [n_inv][order]
    int32 type=2, 2-D float64 coefficients       [n_inv][n_terms]
    int32 type=3, 3-D int32   m-combinations     [n_inv][n_terms][order]

This script reads them as data and writes one compressed ``.npz`` with, for
each order o, the arrays ``l{o}`` (n_inv x o), ``nt{o}`` (n_inv),
``m{o}`` (sum n_terms x o) and ``c{o}`` (sum n_terms). It is run once, in a
container where the reference tree is present; the package reads the
``.npz``.

Usage: python convert_gtinv_tables.py [REF_POLYMLP_DIR] [OUT.npz]
"""

import os
import struct
import sys

import numpy as np

DEFAULT_SRC = "/root/reference/src/pypolymlp/cxx/src/polymlp"


def _i32(buf, off):
    return struct.unpack_from("<i", buf, off)[0], off + 4


def _read_2d(buf, off, dtype, size):
    n, off = _i32(buf, off)
    rows = []
    for _ in range(n):
        m, off = _i32(buf, off)
        rows.append(np.frombuffer(buf, dtype=dtype, count=m, offset=off).copy())
        off += m * size
    return rows, off


def _read_3d(buf, off):
    a, off = _i32(buf, off)
    mats = []
    for _ in range(a):
        b, off = _i32(buf, off)
        rows = []
        for _ in range(b):
            c, off = _i32(buf, off)
            rows.append(np.frombuffer(buf, dtype="<i4", count=c, offset=off).copy())
            off += 4 * c
        mats.append(rows)
    return mats, off


def read_table(path):
    """Return (l_combs, coeffs, m_combs) of one table file."""
    buf = open(path, "rb").read()
    if buf[:4] != b"DATA":
        raise ValueError(f"{path}: bad magic")
    off = 4
    _, off = _i32(buf, off)  # number of blocks
    _, off = _i32(buf, off)  # block type 1
    l_all, off = _read_2d(buf, off, "<i4", 4)
    _, off = _i32(buf, off)  # block type 2
    c_all, off = _read_2d(buf, off, "<f8", 8)
    _, off = _i32(buf, off)  # block type 3
    m_all, off = _read_3d(buf, off)
    return l_all, c_all, m_all


def main(src=DEFAULT_SRC, out=None):
    if out is None:
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gtinv_v1.npz")
    arrays = {}
    for order in range(1, 7):
        l_all, c_all, m_all = read_table(
            os.path.join(src, f"polymlp_gtinv_data_v1_order{order}.bin")
        )
        arrays[f"l{order}"] = np.array(l_all, dtype=np.int16).reshape(-1, order)
        arrays[f"nt{order}"] = np.array([len(c) for c in c_all], dtype=np.int32)
        arrays[f"c{order}"] = np.concatenate(c_all).astype(np.float64)
        arrays[f"m{order}"] = (
            np.concatenate([np.array(m, dtype=np.int8).reshape(-1, order) for m in m_all])
        )
    np.savez_compressed(out, **arrays)
    print("wrote", out)


if __name__ == "__main__":
    main(*sys.argv[1:])
