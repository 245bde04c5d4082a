"""Host logic of the refinement step (mlp_dev.refine) with the device calls
replaced by numpy on a small dense problem: rows X (weighted), targets y,
the scaled exact system A = D^-1 X^T X D^-1, and a perturbed, still positive
definite A~ standing in for the quantised Gram.

The ridge solution of A~ is off by O(|A~ - A| / alpha); one refinement step
with the exact residual must shrink that error by about the same factor, and
the squared error it reports must equal |y - X c|^2 of the corrected
coefficients up to the second-order term it takes from A~."""

import numpy as np

from pypolymlp_amd import mlp_dev


class _FakeAcc:
    def __init__(self, X, y, A_tilde, scales, zero):
        self.X, self.y, self.At, self.s, self.zero = X, y, A_tilde, scales, zero

    def residual(self, batch, w, yw, C, precision):
        U = self.y[:, None] - self.X @ C
        return self.X.T @ U, np.sum(U * U, axis=0)

    def ridge_correct(self, scales, zero, alphas, rhs):
        keep = ~zero
        d = np.zeros_like(rhs)
        dad = np.zeros(len(alphas))
        for k, a in enumerate(alphas):
            A = self.At[np.ix_(keep, keep)] + a * np.eye(keep.sum())
            d[keep, k] = np.linalg.solve(A, rhs[keep, k])
            dad[k] = d[keep, k] @ self.At[np.ix_(keep, keep)] @ d[keep, k]
        return d, np.zeros(len(alphas), dtype=np.int32), dad


def test_refine_step_on_perturbed_system():
    rng = np.random.default_rng(0)
    n, F = 400, 40
    X = rng.normal(size=(n, F)) @ np.diag(np.logspace(0, -3, F)) * rng.uniform(0.5, 2, size=F)
    y = X @ rng.normal(size=F) + 1e-2 * rng.normal(size=n)
    s = np.sqrt(np.mean(X * X, axis=0))
    zero = np.zeros(F, dtype=bool)
    zero[5] = True  # a dropped column: zero coefficient, zero correction
    s[5] = 1.0
    keep = ~zero
    Xs = X / s
    A = Xs.T @ Xs
    b = Xs.T @ y
    E = rng.normal(size=(F, F)) * 1e-7 * np.abs(A).max()
    A_t = A + 0.5 * (E + E.T)
    alphas = np.array([1e-4, 1e-2])

    def solve(M, a):
        c = np.zeros(F)
        c[keep] = np.linalg.solve(M[np.ix_(keep, keep)] + a * np.eye(keep.sum()), b[keep])
        return c

    exact = np.stack([solve(A, a) for a in alphas], 1)
    approx = np.stack([solve(A_t, a) for a in alphas], 1)
    assert np.linalg.eigvalsh(A_t[np.ix_(keep, keep)]).min() > -alphas[0]
    dn = mlp_dev.DeviceNormal(acc=_FakeAcc(X, y, A_t, s, zero), scales=s, zero=zero,
                              min_energy=0.0, total_n_data=n, y_sq_norm=float(y @ y),
                              parts=[(None, None, None)])
    ref, sq, info = mlp_dev.refine(dn, approx, alphas, 0)
    assert np.all(info == 0)
    assert np.all(ref[zero] == 0.0)
    for k in range(len(alphas)):
        e0 = np.linalg.norm(approx[:, k] - exact[:, k])
        e1 = np.linalg.norm(ref[:, k] - exact[:, k])
        assert e0 > 0 and e1 < 1e-2 * e0, (k, e0, e1)
        u = y - X @ (ref[:, k] / s)
        assert abs(sq[k] - u @ u) <= 1e-6 * (u @ u)
