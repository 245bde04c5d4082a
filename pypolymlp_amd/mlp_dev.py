"""Regression on the GPU: weights, fused X^T X / X^T y, scales, ridge, RMSE.

Host-side mirror of pypolymlp's sequential regression:
  calc_xtx_xty                 src/pypolymlp/mlp_dev/core/data_sequential.py:25-94
  _compute_products_single_batch                                     :97-155
  apply_weights and helpers    src/pypolymlp/mlp_dev/core/utils_weights.py:10-98
  compute_scales / round_scales src/pypolymlp/mlp_dev/core/utils_scales.py:6-40
  solver_ridge                 src/pypolymlp/mlp_dev/standard/solvers.py:48-84
  compute_mse / compute_rmse   src/pypolymlp/mlp_dev/core/utils_model_selection.py:37-89
  fit                          src/pypolymlp/mlp_dev/standard/fit.py:12-63

The per-row weights and weighted targets are computed here exactly as the
reference computes them; X, X^T X, X^T y and the energy-row sums come from
the HIP library (pypolymlp_amd.capi); the ridge solve runs on the GPU
(plm_ridge: FP64 Cholesky with one refinement step). There is no CPU path:
without the HIP library every entry point raises.
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .datasets import Dataset


# --- weights (utils_weights.py) ---------------------------------------------
def _weight_energy(energy, total_n_atoms, min_e=None):
    e_per_atom = energy / total_n_atoms
    if min_e is None:
        min_e = np.min(e_per_atom)
    e_th1, e_th2 = min_e * 0.75, min_e * 0.50
    w = np.ones(len(energy))
    w[e_per_atom > e_th1] = 0.5
    w[e_per_atom > e_th2] = 0.3
    w[e_per_atom > 0.0] = 0.1
    return w


def _weight_force(forces, tol=1e-12):
    w = np.abs(forces)
    w[w < tol] = tol
    w = np.reciprocal(w)
    w[w > 1.0] = 1.0
    return w


def _weight_stress(stress, weight_stress, tol=1e-12):
    nonzero = np.abs(stress) > tol
    log1 = np.ones(len(stress)) * np.log10(tol)
    log1[nonzero] = np.log10(np.abs(stress)[nonzero])
    w = np.power(5, -log1)
    w[w > 1.0] = 1.0
    w *= weight_stress
    return w


def row_weights(ds: Dataset, min_e: float, weight_stress: float = 0.1):
    """(w, y_w) over the rows of one dataset batch in the reference layout
    [energies | stresses (6 per structure) | forces (3N per structure)]."""
    ne = len(ds.energies)
    n_s = 6 * ne
    forces = ds.forces_flat
    stresses = ds.stresses_flat
    n = ne + n_s + len(forces)
    w = np.ones(n)
    y = np.zeros(n)
    we = _weight_energy(ds.energies, ds.total_n_atoms, min_e=min_e) * ds.weight
    w[:ne] = we
    y[:ne] = we * ds.energies
    if ds.include_force:
        wf = _weight_force(forces) * ds.weight
        w[ne + n_s:] = wf
        y[ne + n_s:] = wf * forces
    else:
        # energy-only dataset: the reference builds no force or stress rows for
        # it (Features, features.py:94-133), so they carry no weight here
        w[ne:] = 0.0
        y[ne:] = 0.0
        return w, y
    if ds.include_stress:
        ws = _weight_stress(stresses, weight_stress * ds.weight)
        w[ne:ne + n_s] = ws
        y[ne:ne + n_s] = ws * stresses
    else:
        w[ne:ne + n_s] = 0.0
        y[ne:ne + n_s] = 0.0
    return w, y


def reference_n_rows(ds: Dataset, n_structures: Optional[int] = None) -> int:
    """Rows the reference's X holds for the first n_structures of ds: energy
    rows, plus 6 stress and 3N force rows per structure when the dataset has
    forces (PyModel::set_index, py_model.cpp:58-106); the RMSE denominator
    total_n_data of data_sequential.py:155."""
    n = len(ds.structures) if n_structures is None else n_structures
    if not ds.include_force:
        return n
    return n + 6 * n + 3 * int(ds.total_n_atoms[:n].sum())


# --- scales (utils_scales.py) ------------------------------------------------
def round_scales(scales, include_force=True, threshold=1e-10):
    zero = np.abs(scales) < (threshold if include_force else threshold * threshold)
    scales[zero] = 1.0
    return scales, zero


def compute_scales(scales, xe_sum, xe_sq_sum, n_data, include_force=True, threshold=1e-10):
    if scales is None:
        var = xe_sq_sum / n_data - np.square(xe_sum / n_data)
        var[var < 0.0] = 1.0
        scales = np.sqrt(var)
    return round_scales(scales, include_force=include_force, threshold=threshold)


def get_min_energy(datasets: Sequence[Dataset]) -> float:
    return float(min(np.min(d.energies / d.total_n_atoms) for d in datasets))


@dataclass
class NormalEquations:
    xtx: np.ndarray
    xty: np.ndarray
    y_sq_norm: float
    total_n_data: int
    scales: np.ndarray
    min_energy: float
    xe_sum: Optional[np.ndarray] = None
    xe_sq_sum: Optional[np.ndarray] = None


def calc_xtx_xty(model: capi.Model, datasets: Sequence[Dataset], precision=capi.PLM_FP64,
                 scales: Optional[np.ndarray] = None, min_energy: Optional[float] = None,
                 weight_stress: float = 0.1, batch_size: int = 4096,
                 scale_threshold: float = 1e-10) -> NormalEquations:
    """Fused X -> (X^T X, X^T y) on the GPU over datasets in structure batches."""
    if min_energy is None:
        min_energy = get_min_energy(datasets)
    acc = capi.Normal(model)
    for ds in datasets:
        ds.sort_by_atoms()
        _accumulate(model, acc, ds, 0, len(ds.structures), min_energy, weight_stress,
                    batch_size, precision)
    n_data = sum(len(d.energies) for d in datasets)
    return _finish_normal(acc.result(), n_data, scales, datasets[0].include_force,
                          scale_threshold, min_energy,
                          n_rows=sum(reference_n_rows(d) for d in datasets))


def _accumulate(model, acc, ds: Dataset, i0: int, i1: int, min_energy, weight_stress,
                batch_size, precision):
    """Add structures [i0, i1) of ds to the device accumulator in batches."""
    for b0 in range(i0, i1, batch_size):
        sub = ds.subset(slice(b0, min(i1, b0 + batch_size)))
        w, yw = row_weights(sub, min_energy, weight_stress)
        sts = sub.structures
        batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                           [s.types for s in sts])
        acc.accumulate(batch, w, yw, precision)


def _finish_normal(res, n_data, scales, include_force, scale_threshold, min_energy,
                   n_rows: int) -> NormalEquations:
    """Scale the raw sums: zero-scale columns dropped, xtx / (s s^T), xty / s.
    n_rows: the reference's row count of the accumulated structures (the
    device batch also carries force rows of energy-only datasets, at weight 0)."""
    xtx, xty, xe_sum, xe_sq, y_sq, _ = res
    scales, zero = compute_scales(None if scales is None else np.array(scales, dtype=float),
                                  xe_sum, xe_sq, n_data, include_force=include_force,
                                  threshold=scale_threshold)
    xtx[zero] = 0.0
    xtx[:, zero] = 0.0
    xty[zero] = 0.0
    xtx /= scales[:, None]
    xtx /= scales[None, :]
    xty /= scales
    return NormalEquations(xtx=xtx, xty=xty, y_sq_norm=y_sq, total_n_data=n_rows, scales=scales,
                           min_energy=min_energy, xe_sum=xe_sum, xe_sq_sum=xe_sq)


# --- the standard path: X in memory (features.py, data_standard.py) ----------
@dataclass
class Features:
    x: np.ndarray
    first_indices: list  # (ebegin, fbegin, sbegin) per dataset, -1 without forces
    n_data: tuple        # (n energy, n force, n stress) rows
    cumulative_n_features: Optional[list] = None  # hybrid models: end column of each sub-model


def compute_features(model: capi.Model, datasets: Sequence[Dataset],
                     precision=capi.PLM_FP64) -> Features:
    """X of all datasets in the reference layout (Features, features.py:94-133;
    PyModel::set_index, src/cxx/src/compute/py_model.cpp:58-104): energy rows
    of every structure, then 6 stress rows and then 3N force rows of the
    structures of datasets with forces. X is built on the GPU over one batch
    of all structures; rows of datasets without forces are dropped."""
    sts = [s for d in datasets for s in d.structures]
    batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                       [s.types for s in sts])
    X = capi.features_x(model, batch, precision)
    n_st = len(sts)
    n_at = np.array([s.n_total for s in sts])
    f_off = n_st * 7 + np.concatenate([[0], np.cumsum(3 * n_at)])
    e_rows, s_rows, f_rows, first = [np.arange(n_st)], [], [], []
    i0 = 0
    ns = n_st
    nf = n_st + 6 * sum(len(d.structures) for d in datasets if d.include_force)
    for d in datasets:
        n = len(d.structures)
        if d.include_force:
            first.append((i0, nf, ns))
            s_rows.append(n_st + 6 * i0 + np.arange(6 * n))
            f_rows.append(np.arange(f_off[i0], f_off[i0 + n]))
            ns += 6 * n
            nf += 3 * int(n_at[i0:i0 + n].sum())
        else:
            first.append((i0, -1, -1))
        i0 += n
    rows = np.concatenate(e_rows + s_rows + f_rows)
    n_e, n_s = n_st, sum(len(r) for r in s_rows)
    return Features(x=X[rows], first_indices=first,
                    n_data=(n_e, len(rows) - n_e - n_s, n_s))


@dataclass
class DataXY:
    x: np.ndarray
    y: np.ndarray
    weights: np.ndarray
    scales: np.ndarray
    min_energy: float
    first_indices: list
    n_structures: int
    cumulative_n_features: Optional[list] = None


def calc_xy(model: capi.Model, datasets: Sequence[Dataset], scales=None, min_energy=None,
            weight_stress: float = 0.1, scale_threshold: float = 1e-10,
            precision=capi.PLM_FP64) -> DataXY:
    """Weighted, scaled X and y in memory (calc_xy, data_standard.py:16-77):
    scales = population std of the energy rows, columns with zero scale are
    zeroed, rows weighted as apply_weights (utils_weights.py:50-98)."""
    feat = compute_features(model, datasets, precision)
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
    return DataXY(x=x, y=y, weights=w, scales=scales, min_energy=float(min_energy),
                  first_indices=feat.first_indices, n_structures=ne)


def solver_ridge(xtx, xty, alphas):
    """Coefficients (F x n_alpha) of (xtx + alpha I) c = xty for each alpha, on
    the GPU (solvers.py:48-84). A failed factorization gives 1e30 entries, as
    solvers.py:41-43 does."""
    return capi.ridge(xtx, xty, alphas).T


def compute_rmse(coefs_array, ne: NormalEquations):
    out = []
    for c in coefs_array.T:
        mse = (c @ (ne.xtx @ c) - 2 * c @ ne.xty + ne.y_sq_norm) / ne.total_n_data
        out.append(np.sqrt(mse) if mse >= 0 else 1e10)
    return np.array(out)


@dataclass
class FitResult:
    coeffs: np.ndarray
    scales: np.ndarray
    alpha: float
    rmse_train: float
    rmse_test: float

    @property
    def coeffs_unscaled(self):
        """Coefficients as stored in polymlp.yaml and passed to evaluation."""
        return self.coeffs / self.scales


@dataclass
class DeviceNormal:
    """Normal equations left on the GPU: the accumulator (X^T X, X^T y, y^T y
    of [X | y]) and what the host needs from it. parts: the device batches
    with their row weights and weighted targets, when kept for a residual
    pass over the same rows."""
    acc: capi.Normal
    scales: np.ndarray
    zero: np.ndarray
    min_energy: float
    total_n_data: int
    y_sq_norm: float
    parts: Optional[list] = None


def _parts(model, datasets: Sequence[Dataset], min_energy, weight_stress, batch_size):
    """(device batch, w, y_w) per batch of batch_size structures, datasets
    sorted by atom count (the batching of _accumulate)."""
    for ds in datasets:
        ds.sort_by_atoms()
        for b0 in range(0, len(ds.structures), batch_size):
            sub = ds.subset(slice(b0, min(len(ds.structures), b0 + batch_size)))
            w, yw = row_weights(sub, min_energy, weight_stress)
            sts = sub.structures
            batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                               [s.types for s in sts])
            yield batch, w, yw


def accumulate_normal(model: capi.Model, datasets: Sequence[Dataset], precision=capi.PLM_FP64,
                      scales: Optional[np.ndarray] = None, min_energy: Optional[float] = None,
                      weight_stress: float = 0.1, batch_size: int = 4096,
                      scale_threshold: float = 1e-10, keep_parts: bool = False) -> DeviceNormal:
    """calc_xtx_xty (data_sequential.py:25-94) without reading X^T X back: the
    accumulation is the same; only the energy-row sums and y^T y come to the
    host, where the scales are computed (utils_scales.py:6-40). The scaling
    of X^T X and X^T y is applied on the device by DeviceNormal.acc.ridge /
    sq_error. keep_parts: keep the device batches for a residual pass, and
    their weighted rows on the device up to PLM_KEEP_ROWS_GB (default 64) GiB,
    so that the pass reads them instead of building them again."""
    if min_energy is None:
        min_energy = get_min_energy(datasets)
    acc = capi.Normal(model)
    kept = [] if keep_parts else None
    if keep_parts:
        acc.keep_rows(int(float(os.environ.get("PLM_KEEP_ROWS_GB", "64")) * (1 << 30)))
    for batch, w, yw in _parts(model, datasets, min_energy, weight_stress, batch_size):
        acc.accumulate(batch, w, yw, precision)
        if keep_parts:
            kept.append((batch, w, yw))
    xe_sum, xe_sq, y_sq, _ = acc.vectors()
    n_data = sum(len(d.energies) for d in datasets)
    scales, zero = compute_scales(None if scales is None else np.array(scales, dtype=float),
                                  xe_sum, xe_sq, n_data, include_force=datasets[0].include_force,
                                  threshold=scale_threshold)
    return DeviceNormal(acc=acc, scales=scales, zero=zero, min_energy=float(min_energy),
                        total_n_data=sum(reference_n_rows(d) for d in datasets), y_sq_norm=y_sq,
                        parts=kept)


def residual_sums(acc: capi.Normal, parts, coeffs_unscaled, precision=capi.PLM_FP64):
    """(X^T (y - X c) (F, k), |y - X c|^2 (k)) over the rows of parts for each
    column c of coeffs_unscaled (F, k) (plm_normal_residual per batch)."""
    C = np.asarray(coeffs_unscaled, dtype=float)
    R = np.zeros(C.shape)
    sq = np.zeros(C.shape[1])
    for batch, w, yw in parts:
        r, s = acc.residual(batch, w, yw, C, precision)
        R += r
        sq += s
    return R, sq


def refine(dn: DeviceNormal, coefs, alphas, precision):
    """One step of iterative refinement of ridge solutions (F, n_alpha) of
    the scaled system of dn, whose Gram is the quantised one: the residual of
    the normal equations of X itself, D^-1 X^T (y - X D^-1 c) - alpha c, from a
    residual pass over the kept rows, then the correction solved with the
    accumulated system. Returns (coefs, squared errors of the corrected
    solutions, info). The squared error is |y - X c|^2 of the pass minus
    2 delta^T g plus delta^T A delta (the last term from the accumulated
    system, second order in the correction)."""
    alphas = np.asarray(alphas, dtype=float)
    sc = dn.scales[:, None]
    R, sq_u = residual_sums(dn.acc, dn.parts, coefs / sc, precision)
    g = R / sc
    g[dn.zero] = 0.0
    rhs = g - alphas[None, :] * coefs
    delta, info, dad = dn.acc.ridge_correct(dn.scales, dn.zero, alphas, rhs)
    delta[:, info != 0] = 0.0
    return coefs + delta, sq_u - 2.0 * np.sum(delta * g, axis=0) + dad, info


def _rmse_from_sq(sq, n_data):
    """compute_rmse (utils_model_selection.py:37-47) from the squared-error sums."""
    mse = np.asarray(sq, dtype=float) / n_data
    return np.where(mse >= 0, np.sqrt(np.abs(mse)), 1e10)


def fit(model: capi.Model, train: List[Dataset], test: List[Dataset], alphas,
        precision=capi.PLM_FP64, refine_quantised: bool = True, weight_stress: float = 0.1,
        batch_size: int = 4096) -> FitResult:
    """Sequential ridge fit with model selection on the test RMSE (fit.py:12-63).
    The normal equations stay on the GPU: they are scaled and solved there
    (plm_normal_ridge); only vectors cross to the host. With the quantised
    Gram (PLM_MIXED, PLM_MIXED64) each solution gets one step of iterative
    refinement against the rows themselves (refine). The training RMSE comes
    from the device system (plm_normal_sq_error's form) or, after refinement,
    from the residual pass; the test RMSE from a residual pass over the test
    rows with the X build of the fit (no test Gram is formed)."""
    alphas = np.asarray(alphas, dtype=float)
    quantised = precision != capi.PLM_FP64
    do_refine = quantised and refine_quantised
    tr = accumulate_normal(model, train, precision=precision, weight_stress=weight_stress,
                           batch_size=batch_size, keep_parts=do_refine)
    coefs, info, sq = tr.acc.ridge(tr.scales, tr.zero, alphas)
    ok = info == 0
    coefs[:, ~ok] = 0.0
    if do_refine:
        coefs, sq, info2 = refine(tr, coefs, alphas, precision)
        ok &= info2 == 0
        tr.parts = None
        tr.acc.keep_rows(0)
    sq = np.where(ok, sq, np.inf)
    rmse_train = _rmse_from_sq(sq, tr.total_n_data)
    te_parts = _parts(model, test, tr.min_energy, weight_stress, batch_size)
    _, sq_te = residual_sums(tr.acc, te_parts, coefs / tr.scales[:, None], precision)
    sq_te = np.where(ok, sq_te, np.inf)
    rmse_test = _rmse_from_sq(sq_te, sum(reference_n_rows(d) for d in test))
    coefs[:, ~ok] = 1e30
    k = int(np.argmin(rmse_test))
    return FitResult(coeffs=coefs[:, k], scales=tr.scales, alpha=float(alphas[k]),
                     rmse_train=float(rmse_train[k]), rmse_test=float(rmse_test[k]))


# --- conjugate gradient (gradient/solvers_cg.py, gradient/fit_cg.py) --------
def solver_cg(xtx, xty, alpha: float, coef0=None, gtol: float = 1e-2, max_iter: int = 10000):
    """CG on (xtx + alpha I) c = xty from coef0 (zeros), on the GPU
    (solver_cg, src/pypolymlp/mlp_dev/gradient/solvers_cg.py:32-93): the loop
    ends at the first iteration i > 100 with |r| / sqrt(F) < gtol or after
    max_iter iterations, and returns the coefficients it ends with (the
    reference's coef_min aliases the iterate, which is updated in place)."""
    return capi.solver_cg(xtx, xty, alpha, coef0=coef0, gtol=gtol, max_iter=max_iter)[0]


def fit_cg(model: capi.Model, train: List[Dataset], test: List[Dataset], alphas,
           precision=capi.PLM_FP64, gtol: float = 1e-2, max_iter: Optional[int] = None
           ) -> FitResult:
    """fit_cg (src/pypolymlp/mlp_dev/gradient/fit_cg.py:21-80): CG solves from the
    largest alpha down, each warm-started from the previous solution, then the
    alpha with the smallest test RMSE."""
    tr = calc_xtx_xty(model, train, precision=precision)
    if max_iter is None:
        max_iter = max(model.n_features * 3, 50000)
    alphas = np.asarray(alphas, dtype=float)
    coefs, c0 = [], None
    for a in alphas[::-1]:
        c0 = solver_cg(tr.xtx, tr.xty, a, coef0=c0, gtol=gtol, max_iter=max_iter)
        coefs.append(c0)
    coefs = np.array(coefs)[::-1].T
    rmse_train = compute_rmse(coefs, tr)
    te = calc_xtx_xty(model, test, precision=precision, scales=tr.scales.copy(),
                      min_energy=tr.min_energy)
    rmse_test = compute_rmse(coefs, te)
    k = int(np.argmin(rmse_test))
    return FitResult(coeffs=coefs[:, k], scales=tr.scales, alpha=float(alphas[k]),
                     rmse_train=float(rmse_train[k]), rmse_test=float(rmse_test[k]))


# --- accuracy and learning curve (eval_accuracy.py, fit.py:115-179) ---------
def _rmse(true, pred):
    return float(np.sqrt(np.mean(np.square(true - pred))))


def eval_accuracy(model: capi.Model, coeffs_unscaled, datasets: Sequence[Dataset],
                  precision=capi.PLM_FP64) -> dict:
    """RMSEs of the potential with ``coeffs_unscaled`` (coeffs / scales) over each
    dataset: energy per atom (eV/atom), force (eV/A), stress per atom (eV/atom,
    None without stress data), evaluated on the GPU
    (PolymlpEvalAccuracy.compute_error_single, eval_accuracy.py:49-150)."""
    model.set_coeffs(np.asarray(coeffs_unscaled, dtype=float))
    out = {}
    for ds in datasets:
        sts = ds.structures
        batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                           [s.types for s in sts])
        E, F, S = capi.eval_batch(model, batch, precision)
        n = ds.total_n_atoms
        err = {"energy": _rmse(ds.energies / n, E / n), "force": None, "stress": None}
        if ds.exist_force:
            err["force"] = _rmse(ds.forces_flat, np.concatenate([f.reshape(-1) for f in F]))
        if ds.exist_stress:
            nn = np.repeat(n, 6)
            err["stress"] = _rmse(ds.stresses_flat / nn, S.reshape(-1) / nn)
        out[ds.name] = err
    return out


def fit_learning_curve(model: capi.Model, train: List[Dataset], test: List[Dataset], alphas,
                       precision=capi.PLM_FP64, weight_stress: float = 0.1,
                       scale_threshold: float = 1e-10):
    """Learning curve (fit_learning_curve, fit.py:115-179): for n = n_train/10,
    2 n_train/10, ... training structures (in dataset order), ridge fits of the
    first n structures with the scales and minimum energy of the whole training
    set, the alpha chosen by test RMSE, and the test errors of that model.

    One pass builds X once: the device accumulator's sums are read after each
    prefix, so prefix k costs only its own n_train/10 structures.
    Returns [[n, errors], ...], one entry per (n, test dataset)."""
    if len(train) != 1:
        raise RuntimeError("Number of training datasets must be one for learning curve.")
    ds = train[0]
    n_train = len(ds.structures)
    step = n_train // 10
    if step == 0:
        raise ValueError("learning curve needs at least 10 training structures")
    min_e = get_min_energy(train)
    acc = capi.Normal(model)
    prefixes, done = [], 0
    for n in range(step, n_train + 1, step):
        _accumulate(model, acc, ds, done, n, min_e, weight_stress, 4096, precision)
        done = n
        prefixes.append((n, acc.result()))
    _accumulate(model, acc, ds, done, n_train, min_e, weight_stress, 4096, precision)
    full = _finish_normal(acc.result(), n_train, None, ds.include_force, scale_threshold, min_e,
                          n_rows=reference_n_rows(ds))
    scales = full.scales
    te = calc_xtx_xty(model, test, precision=precision, scales=scales.copy(), min_energy=min_e,
                      weight_stress=weight_stress, scale_threshold=scale_threshold)
    log = []
    for n, res in prefixes:
        ne = _finish_normal(res, n, scales, ds.include_force, scale_threshold, min_e,
                            n_rows=reference_n_rows(ds, n))
        coefs = solver_ridge(ne.xtx, ne.xty, alphas)
        rmse_train = compute_rmse(coefs, ne)
        rmse_test = compute_rmse(coefs, te)
        k = int(np.argmin(rmse_test))
        errors = eval_accuracy(model, coefs[:, k] / scales, test, precision=capi.PLM_FP64)
        for v in errors.values():
            log.append([n, v])
    return log
