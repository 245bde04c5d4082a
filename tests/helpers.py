"""Shared helpers for the tests (structures, models, fixtures)."""

from __future__ import annotations

import os

import numpy as np

from pypolymlp_amd.params import load_mlp_yaml
from pypolymlp_amd.structure import read_poscar

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, "golden", "ref_files")


def ref_file(name: str) -> str:
    return os.path.join(FILES, name)


def load_model(name: str):
    """(ModelSpec, coeffs) of a model file shipped with the reference tests."""
    return load_mlp_yaml(ref_file(name))


def poscar(name: str):
    return read_poscar(ref_file(name))


def random_cell_structure(rng, n_atoms=6, n_type=2, spread=(-0.3, 1.3), skew=1.5):
    """Skewed cell with fractional coordinates partly outside [0, 1)."""
    while True:
        ax = np.eye(3) * rng.uniform(4.0, 7.0) + rng.normal(0.0, skew, (3, 3))
        if abs(np.linalg.det(ax)) > 30.0:
            break
    frac = rng.uniform(spread[0], spread[1], (3, n_atoms))
    types = rng.integers(0, n_type, n_atoms)
    return ax, ax @ frac, types


def write_model_yaml(spec, path, coeffs=None):
    from pypolymlp_amd.params import save_mlp_yaml
    from pypolymlp_amd import capi

    if coeffs is None:
        n = capi.Model(spec.as_dict()).n_features
        coeffs = np.ones(n)
    save_mlp_yaml(spec, coeffs, np.ones(len(coeffs)), path)
    return path


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def oracle_x_and_noise(R, axes, pos, types, n_rot=2, seed=0):
    """Oracle X rows and, per element, the oracle's own rounding spread.

    The spread is measured under exact invariances: rigid rotations of the
    whole structure (energy rows invariant, force rows rotate as vectors,
    stress rows as a symmetric tensor) mapped back to the original frame.
    Features formed by heavy cancellation (near-symmetric cells) are fixed
    by the reference only to this spread.
    """
    Xr = R.x_batch(axes, pos, types, force=True)
    n_st = len(axes)
    n_at = [len(t) for t in types]
    noise = np.zeros_like(Xr)
    rng = np.random.default_rng(seed)
    for _ in range(n_rot):
        Q = _rotation(rng)
        Xq = R.x_batch([Q @ a for a in axes], [Q @ p for p in pos], types, force=True)
        back = Xq.copy()
        srow = n_st
        frow = n_st + 6 * n_st
        for s in range(n_st):
            S = Xq[srow + 6 * s: srow + 6 * s + 6]  # xx yy zz xy yz zx
            T = np.array([[S[0], S[3], S[5]], [S[3], S[1], S[4]], [S[5], S[4], S[2]]])
            U = np.einsum("ai,abc,bj->ijc", Q, T, Q)
            back[srow + 6 * s: srow + 6 * s + 6] = [U[0, 0], U[1, 1], U[2, 2], U[0, 1], U[1, 2], U[2, 0]]
            Fq = Xq[frow: frow + 3 * n_at[s]].reshape(n_at[s], 3, -1)
            back[frow: frow + 3 * n_at[s]] = np.einsum("ai,kac->kic", Q, Fq).reshape(3 * n_at[s], -1)
            frow += 3 * n_at[s]
        noise = np.maximum(noise, np.abs(back - Xr))
    return Xr, noise


def x_parity(X, Xr, noise, resolved=1e-12):
    """Column-relative errors split by how well the reference resolves a column.

    Returns (strict, loose, n_loose): the max column-relative error
    max|X - Xr| / max|Xr| over columns whose oracle rotation spread is at most
    `resolved` of the column max, the same over the remaining
    (cancellation-dominated) columns, and their number.
    """
    colmax = np.abs(Xr).max(axis=0)
    colmax[colmax == 0] = 1.0
    err = np.abs(X - Xr).max(axis=0)
    rel = err / colmax
    spread = noise.max(axis=0) / colmax
    ok = spread <= resolved
    strict = float(rel[ok].max()) if ok.any() else 0.0
    # unresolved columns: relative to max(column max, 1e-9 * median column max),
    # so columns that are numerically zero (~1e-15 in near-symmetric cells)
    # are compared on the scale of the typical feature
    floor = np.maximum(colmax, 1e-9 * np.median(colmax))
    loose = float((err / floor)[~ok].max()) if (~ok).any() else 0.0
    return strict, loose, int((~ok).sum())


# --- the mixed Gram's quantisation, restated in numpy
# (pypolymlp_amd/csrc/kernels_gram.hip: colexp_kernel, quantize_kernel)
QUANT_CHUNK = 4096


def quant_digits(x, e):
    """Three signed digits of x at exponent e, x~ = 2^(e-7) (d0 + d1/128 + d2/16384)."""
    m = np.ldexp(x, 7 - e)
    t0 = np.trunc(m)
    r1 = (m - t0) * 128
    t1 = np.trunc(r1)
    t2 = np.clip(np.rint((r1 - t1) * 128), -127, 127)
    return t0, t1, t2


def quantised_gram(A, bounds, max_level=4):
    """Gram of A quantised per (row chunk, column): the smallest exponent e with
    |x| < 2^e over the chunk, three digits each; the sum of the digit products
    d_s^T d_t with s + t <= max_level (4: all nine, the exact Gram of the
    quantised matrix). Each product is an exact integer in FP64."""
    F1 = A.shape[1]
    G = np.zeros((F1, F1))
    for a, b in bounds:
        blk = A[a:b]
        m = np.abs(blk).max(axis=0)
        e = np.where(m > 0, np.frexp(m)[1], 0)
        d = quant_digits(blk, e[None, :])
        ee = e[:, None] + e[None, :] - 14
        for s in range(3):
            for t in range(3):
                if s + t <= max_level:
                    G += np.ldexp(d[s].T @ d[t], ee) * 2.0 ** (-7 * (s + t))
    return G


def quant_chunks(n_es, n_rows):
    """Row chunks of a quantisation order that puts the n_es energy and stress
    rows first: one chunk of them, then force-row chunks of QUANT_CHUNK rows."""
    return [(0, n_es)] + [(a, min(a + QUANT_CHUNK, n_rows))
                          for a in range(n_es, n_rows, QUANT_CHUNK)]
