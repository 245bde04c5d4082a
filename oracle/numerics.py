"""Oracle (test infrastructure only): the reference's ridge solve restated.

solver_ridge_posv follows src/pypolymlp/mlp_dev/standard/solvers.py:9-84:
LAPACK posv on the upper triangle for each alpha, with alpha added to the
diagonal cumulatively, and 1e30 coefficients when the factorization fails.
"""

from __future__ import annotations

import numpy as np


def solve_linear_equation(A, b):
    """solvers.py:9-45."""
    from scipy.linalg.lapack import get_lapack_funcs

    (posv,) = get_lapack_funcs(("posv",), (A, b))
    _, x, info = posv(A, b, lower=False, overwrite_a=False, overwrite_b=False)
    if info != 0:
        x = np.ones(x.shape) * 1e30
    return x


def solver_ridge_posv(xtx, xty, alphas):
    """solvers.py:48-84: coefficients (F x n_alpha)."""
    F = xtx.shape[0]
    coefs = np.zeros((F, len(alphas)))
    A = np.array(xtx, dtype=np.float64, copy=True)
    prev = 0.0
    for i, a in enumerate(alphas):
        A.flat[:: F + 1] += a - prev
        coefs[:, i] = solve_linear_equation(A, xty)
        prev = a
    return coefs


def solver_cg(xtx, xty, alpha, coef0=None, gtol=1e-2, max_iter=10000):
    """Conjugate gradient as src/pypolymlp/mlp_dev/gradient/solvers_cg.py:32-93
    restates it (xtx form): (coefficients the loop ends with, iterations).
    The reference's coef_min aliases the in-place iterate, so the final
    iterate is what it returns."""
    F = len(xty)
    coef = np.zeros(F) if coef0 is None else np.array(coef0, dtype=float)
    r = xty - (xtx @ coef + alpha * coef)
    d = r.copy()
    nr = np.linalg.norm(r)
    it = max_iter
    for i in range(max_iter):
        if i > 100 and nr / np.sqrt(F) < gtol:
            it = i
            break
        t = xtx @ d + alpha * d
        lr = nr ** 2 / (t @ d)
        coef += lr * d
        r -= lr * t
        nr_new = np.linalg.norm(r)
        d = r + (nr_new ** 2 / nr ** 2) * d
        nr = nr_new
    return coef, it
