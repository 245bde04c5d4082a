"""ctypes binding of the C-ABI (include/polymlp_amd.h).

The HIP library ``libpolymlp_amd.so`` is built in-tree by ``pypolymlp_amd.build``.
Importing this module does not touch the GPU; device work starts at the first
batch, neighbour-list or evaluation call. A missing library raises at load
time: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpolymlp_amd.so")

PLM_FP64 = 0
PLM_MIXED = 1
PLM_MIXED64 = 2  # FP64 features with the quantised int8 Gram

EXPORTS = [
    "plm_last_error", "plm_device_count", "plm_model_create", "plm_model_destroy",
    "plm_model_n_features", "plm_model_type_sizes", "plm_model_set_coeffs",
    "plm_cell_translations", "plm_neighbor", "plm_batch_create", "plm_batch_destroy",
    "plm_batch_n_rows", "plm_batch_n_pairs", "plm_features_x", "plm_normal_create", "plm_normal_destroy",
    "plm_normal_accumulate", "plm_normal_result", "plm_batch_row_order", "plm_ridge",
    "plm_normal_ridge", "plm_normal_sq_error", "plm_normal_graph_create", "plm_normal_graph_run",
    "plm_normal_graph_destroy", "plm_normal_residual", "plm_normal_ridge_correct",
    "plm_model_tuning", "plm_normal_keep_rows",
    "plm_solver_cg", "plm_model_feature_attrs",
    "plm_eval", "plm_eval_noalloc", "plm_eval_batch", "plm_eval_graph_create", "plm_eval_graph_run",
    "plm_eval_graph_destroy", "plm_profile_enable", "plm_profile_reset", "plm_profile_get",
    "plm_surrogate_n_params", "plm_surrogate_train", "plm_surrogate_grad",
    "plm_surrogate_predict", "plm_get_fn", "plm_get_ylm", "plm_api_features",
    "plm_api_sum_of_prod",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_fp = ctypes.POINTER(ctypes.c_float)


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("n_type", ctypes.c_int), ("feature_type", ctypes.c_int),
        ("model_type", ctypes.c_int), ("max_p", ctypes.c_int), ("max_l", ctypes.c_int),
        ("cutoff", ctypes.c_double), ("n_params", ctypes.c_int), ("params", _dp),
        ("cond_count", _ip), ("cond_index", _ip), ("n_lcomb", ctypes.c_int),
        ("lcomb_order", _ip), ("lcomb_l", _ip), ("lcomb_nterm", _ip), ("lm", _ip),
        ("lm_coeff", _dp),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH} (run python -m pypolymlp_amd.build)"
            )
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.plm_last_error.restype = ctypes.c_char_p
        L.plm_model_create.argtypes = [ctypes.POINTER(ModelDesc), ctypes.POINTER(vp)]
        L.plm_model_destroy.argtypes = [vp]
        L.plm_model_destroy.restype = None
        L.plm_model_n_features.argtypes = [vp, _ip]
        L.plm_model_type_sizes.argtypes = [vp, ctypes.c_int, _ip, _ip, _ip]
        L.plm_model_set_coeffs.argtypes = [vp, _dp, ctypes.c_int]
        L.plm_cell_translations.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, _dp,
                                            ctypes.c_int, _dp, _ip]
        L.plm_neighbor.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   _ip, _ip, _ip, _dp, ctypes.c_long, _lp]
        L.plm_batch_create.argtypes = [vp, ctypes.c_int, _dp, _dp, _ip, _ip, ctypes.POINTER(vp)]
        L.plm_batch_destroy.argtypes = [vp]
        L.plm_batch_destroy.restype = None
        L.plm_batch_n_rows.argtypes = [vp, _lp]
        L.plm_batch_n_pairs.argtypes = [vp, _lp]
        L.plm_batch_row_order.argtypes = [vp, _lp]
        L.plm_features_x.argtypes = [vp, vp, ctypes.c_int, _dp]
        L.plm_normal_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.plm_normal_destroy.argtypes = [vp]
        L.plm_normal_destroy.restype = None
        L.plm_normal_accumulate.argtypes = [vp, vp, _dp, _dp, ctypes.c_int]
        L.plm_normal_result.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _lp]
        L.plm_ridge.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _ip]
        _bp = ctypes.POINTER(ctypes.c_ubyte)
        L.plm_normal_ridge.argtypes = [vp, _dp, _bp, _dp, ctypes.c_int, _dp, _ip, _dp]
        L.plm_normal_sq_error.argtypes = [vp, _dp, _bp, _dp, ctypes.c_int, _dp]
        L.plm_normal_graph_create.argtypes = [vp, vp, _dp, _dp, ctypes.c_int, ctypes.POINTER(vp)]
        L.plm_normal_graph_run.argtypes = [vp, ctypes.c_int]
        L.plm_normal_graph_destroy.argtypes = [vp]
        L.plm_eval.argtypes = [vp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.plm_eval_noalloc.argtypes = [vp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       _dp]
        L.plm_eval_batch.argtypes = [vp, vp, ctypes.c_int, _dp, _dp, _dp]
        L.plm_eval_graph_create.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.plm_eval_graph_run.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp]
        L.plm_eval_graph_destroy.argtypes = [vp]
        L.plm_profile_enable.argtypes = [ctypes.c_int]
        L.plm_surrogate_n_params.argtypes = [_ip]
        L.plm_surrogate_train.argtypes = [_fp, _fp, _fp, ctypes.c_int, _fp, _fp, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_float, ctypes.c_float, _fp]
        L.plm_surrogate_grad.argtypes = [_fp, _fp, _fp, ctypes.c_int, _fp, _fp]
        L.plm_surrogate_predict.argtypes = [_fp, _fp, ctypes.c_int, _fp]
        L.plm_profile_get.argtypes = [ctypes.c_char_p, _dp, _lp]
        L.plm_get_fn.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _dp, _dp]
        L.plm_get_ylm.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.plm_api_features.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.plm_api_sum_of_prod.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().plm_last_error().decode())


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_ip) if a is not None else None


class Model:
    """plm_model built from a params dict (PolymlpParamsSingle.as_dict() schema)."""

    def __init__(self, params_dict: dict):
        m = params_dict["model"]
        n_type = int(params_dict["n_type"])
        keep = []
        d = ModelDesc()
        d.n_type = n_type
        d.feature_type = 1 if m["feature_type"] == "gtinv" else 0
        d.model_type = int(m["model_type"])
        d.max_p = int(m["max_p"])
        d.max_l = int(m["max_l"])
        d.cutoff = float(m["cutoff"])
        params = _d(np.array(m["pair_params"], dtype=float).reshape(-1, 2))
        keep.append(params)
        d.n_params = params.shape[0]
        d.params = _pd(params)
        counts, idx = [], []
        cond = m["pair_params_conditional"]
        for i in range(n_type):
            for j in range(i, n_type):
                v = cond[(i, j)] if (i, j) in cond else cond[str((i, j))]
                counts.append(len(v))
                idx.extend(int(x) for x in v)
        cc, ci = _i(counts), _i(idx if idx else [0])
        keep += [cc, ci]
        d.cond_count, d.cond_index = _pi(cc), _pi(ci)
        if d.feature_type == 1:
            g = m["gtinv"]
            l_comb, lm_seq, coeffs = g["l_comb"], g["lm_seq"], g["lm_coeffs"]
            orders = _i([len(lc) for lc in l_comb])
            l_flat = _i([v for lc in l_comb for v in lc])
            nterm = _i([len(c) for c in coeffs])
            lm_flat = _i([v for seq in lm_seq for row in seq for v in row])
            c_flat = _d([v for c in coeffs for v in c])
            keep += [orders, l_flat, nterm, lm_flat, c_flat]
            d.n_lcomb = len(l_comb)
            d.lcomb_order, d.lcomb_l, d.lcomb_nterm = _pi(orders), _pi(l_flat), _pi(nterm)
            d.lm, d.lm_coeff = _pi(lm_flat), _pd(c_flat)
        h = ctypes.c_void_p()
        _check(lib().plm_model_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.n_type = n_type

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.plm_model_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def n_features(self) -> int:
        n = ctypes.c_int()
        _check(lib().plm_model_n_features(self._h, ctypes.byref(n)))
        return n.value

    def type_sizes(self, t: int):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().plm_model_type_sizes(self._h, t, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_coeffs(self, coeffs):
        c = _d(coeffs)
        _check(lib().plm_model_set_coeffs(self._h, _pd(c), len(c)))


def _pack_structures(axes: Sequence, positions_c: Sequence, types: Sequence):
    axes_a = _d(np.stack([np.asarray(a, dtype=float) for a in axes]))
    pos_a = _d(np.concatenate([_d(p).ravel() for p in positions_c]))
    types_a = _i(np.concatenate([np.asarray(t) for t in types]))
    n_atoms = _i([len(t) for t in types])
    return axes_a, pos_a, types_a, n_atoms


class Batch:
    """Device-resident batch of structures with its full neighbour list."""

    def __init__(self, model: Model, axes, positions_c, types):
        self.model = model
        a, p, t, n = _pack_structures(axes, positions_c, types)
        self.n_atoms = n
        h = ctypes.c_void_p()
        _check(lib().plm_batch_create(model.handle, len(n), _pd(a), _pd(p), _pi(t), _pi(n),
                                      ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.plm_batch_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def n_rows(self) -> int:
        r = ctypes.c_long()
        _check(lib().plm_batch_n_rows(self._h, ctypes.byref(r)))
        return r.value

    @property
    def n_pairs(self) -> int:
        r = ctypes.c_long()
        _check(lib().plm_batch_n_pairs(self._h, ctypes.byref(r)))
        return r.value

    def row_order(self) -> np.ndarray:
        """Batch-internal row index of each row in the reference layout."""
        o = np.zeros(self.n_rows, dtype=np.int64)
        _check(lib().plm_batch_row_order(self._h, o.ctypes.data_as(_lp)))
        return o


def features_x(model: Model, batch: Batch, precision: int = PLM_FP64) -> np.ndarray:
    """X (rows x F) in the reference's PotentialModel layout."""
    X = np.zeros((batch.n_rows, model.n_features))
    _check(lib().plm_features_x(model.handle, batch.handle, precision, _pd(X)))
    return X


class Normal:
    """Device accumulator of X^T X, X^T y, xe_sum and xe_sq_sum."""

    def __init__(self, model: Model):
        self.model = model
        h = ctypes.c_void_p()
        _check(lib().plm_normal_create(model.handle, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.plm_normal_destroy(self._h)
            self._h = None

    def accumulate(self, batch: Batch, w_ref, y_ref, precision: int = PLM_FP64):
        """w_ref / y_ref: weights and weighted targets in the reference row layout."""
        order = batch.row_order()
        w = np.zeros(batch.n_rows)
        y = np.zeros(batch.n_rows)
        w[order] = w_ref
        y[order] = y_ref
        _check(lib().plm_normal_accumulate(self._h, batch.handle, _pd(w), _pd(y), precision))

    def result(self):
        """(xtx, xty, xe_sum, xe_sq_sum, y_sq, n_rows); xty and y_sq come from
        the same Gram pass as xtx."""
        F = self.model.n_features
        xtx = np.zeros((F, F))
        xty = np.zeros(F)
        s = np.zeros(F)
        s2 = np.zeros(F)
        ysq = ctypes.c_double()
        n = ctypes.c_long()
        _check(lib().plm_normal_result(self._h, _pd(xtx), _pd(xty), _pd(s), _pd(s2),
                                       ctypes.byref(ysq), ctypes.byref(n)))
        return xtx, xty, s, s2, ysq.value, n.value

    def vectors(self):
        """(xe_sum, xe_sq_sum, y_sq, n_rows) without reading X^T X back."""
        F = self.model.n_features
        s = np.zeros(F)
        s2 = np.zeros(F)
        ysq = ctypes.c_double()
        n = ctypes.c_long()
        _check(lib().plm_normal_result(self._h, None, None, _pd(s), _pd(s2), ctypes.byref(ysq),
                                       ctypes.byref(n)))
        return s, s2, ysq.value, n.value

    def ridge(self, scales, zero, alphas):
        """(coeffs (F, n_alpha), info, sq_err) of the ridge solves on this
        accumulator's normal equations scaled on the device (plm_normal_ridge)."""
        F = self.model.n_features
        a = _d(alphas)
        sc = _d(scales)
        z = None if zero is None else np.ascontiguousarray(zero, dtype=np.uint8)
        c = np.zeros((len(a), F))
        info = np.zeros(len(a), dtype=np.int32)
        sq = np.zeros(len(a))
        _check(lib().plm_normal_ridge(self._h, _pd(sc), None if z is None else
                                      z.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _pd(a),
                                      len(a), _pd(c), _pi(info), _pd(sq)))
        c[info != 0] = 1e30
        return c.T.copy(), info, sq

    def sq_error(self, scales, zero, coeffs):
        """c^T A c - 2 c^T b + y^T y of each column of coeffs (F, k) on this
        accumulator's normal equations scaled on the device."""
        sc = _d(scales)
        z = None if zero is None else np.ascontiguousarray(zero, dtype=np.uint8)
        C = _d(np.asarray(coeffs).T)
        out = np.zeros(C.shape[0])
        _check(lib().plm_normal_sq_error(self._h, _pd(sc), None if z is None else
                                         z.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _pd(C),
                                         C.shape[0], _pd(out)))
        return out

    def keep_rows(self, max_bytes: int):
        """Keep the weighted rows of later accumulations (up to max_bytes of
        device memory) for residual passes over the same batches
        (plm_normal_keep_rows); 0 releases them."""
        _check(lib().plm_normal_keep_rows(self._h, ctypes.c_ulonglong(int(max_bytes))))

    def residual(self, batch: Batch, w_ref, y_ref, coeffs_unscaled, precision: int = PLM_FP64):
        """(X^T (y - X c) (F, k), |y - X c|^2 (k)) for each column c of
        coeffs_unscaled (F, k), from the weighted rows of batch built as the
        accumulation builds them (plm_normal_residual); up to 8 columns per
        call, more are split."""
        order = batch.row_order()
        w = np.zeros(batch.n_rows)
        y = np.zeros(batch.n_rows)
        w[order] = w_ref
        y[order] = y_ref
        C = _d(np.asarray(coeffs_unscaled, dtype=float).reshape(self.model.n_features, -1).T)
        xtr = np.zeros_like(C)
        sq = np.zeros(C.shape[0])
        for k0 in range(0, C.shape[0], 8):
            k1 = min(C.shape[0], k0 + 8)
            Ck = _d(C[k0:k1])
            part = np.zeros((k1 - k0, C.shape[1]))
            s = np.zeros(k1 - k0)
            _check(lib().plm_normal_residual(self._h, batch.handle, _pd(w), _pd(y), precision,
                                             _pd(Ck), k1 - k0, _pd(part), _pd(s)))
            xtr[k0:k1] = part
            sq[k0:k1] = s
        return xtr.T.copy(), sq

    def ridge_correct(self, scales, zero, alphas, rhs):
        """(delta (F, n_alpha), info, delta^T A delta) of the refinement
        correction (A + alpha_k I) delta_k = rhs_k on this accumulator's scaled
        system (plm_normal_ridge_correct); rhs: (F, n_alpha)."""
        F = self.model.n_features
        a = _d(alphas)
        sc = _d(scales)
        z = None if zero is None else np.ascontiguousarray(zero, dtype=np.uint8)
        R = _d(np.asarray(rhs, dtype=float).reshape(F, len(a)).T)
        d = np.zeros((len(a), F))
        info = np.zeros(len(a), dtype=np.int32)
        dad = np.zeros(len(a))
        _check(lib().plm_normal_ridge_correct(self._h, _pd(sc), None if z is None else
                                              z.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                              _pd(a), len(a), _pd(R), _pd(d), _pi(info), _pd(dad)))
        return d.T.copy(), info, dad


class NormalGraph:
    """One accumulation pass of a batch captured as a HIP graph
    (plm_normal_graph_*): each run(n) adds n passes to the accumulator."""

    def __init__(self, acc: Normal, batch: Batch, w_ref, y_ref, precision: int = PLM_FP64):
        order = batch.row_order()
        w = np.zeros(batch.n_rows)
        y = np.zeros(batch.n_rows)
        w[order] = w_ref
        y[order] = y_ref
        self._keep = (acc, batch)
        h = ctypes.c_void_p()
        _check(lib().plm_normal_graph_create(acc._h, batch.handle, _pd(w), _pd(y), precision,
                                             ctypes.byref(h)))
        self._h = h

    def run(self, n_steps: int = 1):
        _check(lib().plm_normal_graph_run(self._h, int(n_steps)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.plm_normal_graph_destroy(self._h)
            self._h = None


def cell_translations(axis, positions_c, cutoff: float):
    axis, pos = _d(axis), _d(positions_c)
    n = pos.shape[1]
    nt = ctypes.c_int()
    _check(lib().plm_cell_translations(_pd(axis), _pd(pos), n, cutoff, None, 0, None,
                                       ctypes.byref(nt)))
    tr = np.zeros((nt.value, 3))
    prev = np.zeros((3, n))
    _check(lib().plm_cell_translations(_pd(axis), _pd(pos), n, cutoff, _pd(tr), nt.value,
                                       _pd(prev), ctypes.byref(nt)))
    return tr, prev


def neighbor(axis, positions_c, cutoff: float, half: bool = False):
    """(offset[N+1], neigh[M], trans_id[M], d[M,3]) computed on the GPU."""
    axis, pos = _d(axis), _d(positions_c)
    n = pos.shape[1]
    m = ctypes.c_long()
    _check(lib().plm_neighbor(_pd(axis), _pd(pos), n, cutoff, int(half), None, None, None,
                              None, 0, ctypes.byref(m)))
    off = np.zeros(n + 1, dtype=np.int32)
    nb = np.zeros(m.value, dtype=np.int32)
    tid = np.zeros(m.value, dtype=np.int32)
    d = np.zeros((m.value, 3))
    _check(lib().plm_neighbor(_pd(axis), _pd(pos), n, cutoff, int(half), _pi(off), _pi(nb),
                              _pi(tid), _pd(d), m.value, ctypes.byref(m)))
    return off, nb, tid, d


def eval_structure(model: Model, axis, positions_c, types, precision: int = PLM_FP64):
    """(energy, forces (N, 3), stress (6,)) of one structure (PotentialPropertiesFast.eval)."""
    axis, pos = _d(axis), _d(positions_c)
    t = _i(types)
    n = len(t)
    e = np.zeros(1)
    f = np.zeros((n, 3))
    s = np.zeros(6)
    _check(lib().plm_eval(model.handle, _pd(axis), _pd(pos), _pi(t), n, precision, _pd(e), _pd(f),
                          _pd(s)))
    return float(e[0]), f, s


def eval_structure_noalloc(model: Model, axis, positions_c, types, precision: int = PLM_FP64):
    """eval_structure through plm_eval_noalloc: no device allocation in the call
    (the model must already hold its coefficients)."""
    axis, pos = _d(axis), _d(positions_c)
    t = _i(types)
    n = len(t)
    e = np.zeros(1)
    f = np.zeros((n, 3))
    s = np.zeros(6)
    _check(lib().plm_eval_noalloc(model.handle, _pd(axis), _pd(pos), _pi(t), n, precision, _pd(e),
                                  _pd(f), _pd(s)))
    return float(e[0]), f, s


def eval_batch(model: Model, batch: Batch, precision: int = PLM_FP64):
    """(energies (n_st,), forces list of (N_s, 3), stresses (n_st, 6)) of a batch."""
    n_st = len(batch.n_atoms)
    A = int(batch.n_atoms.sum())
    e = np.zeros(n_st)
    f = np.zeros((A, 3))
    s = np.zeros((n_st, 6))
    _check(lib().plm_eval_batch(model.handle, batch.handle, precision, _pd(e), _pd(f), _pd(s)))
    off = np.concatenate([[0], np.cumsum(batch.n_atoms)])
    return e, [f[off[k]:off[k + 1]] for k in range(n_st)], s


class EvalGraph:
    """Evaluation of one batch captured as a HIP graph (plm_eval_graph_*):
    run(n) replays it n times and returns the last replay's results in the
    form of eval_batch. Keep the model and batch alive while it is used; the
    model's coefficients must not change after capture."""

    def __init__(self, model: Model, batch: Batch, precision: int = PLM_FP64):
        self._model, self._batch = model, batch
        h = ctypes.c_void_p()
        _check(lib().plm_eval_graph_create(model.handle, batch.handle, precision, ctypes.byref(h)))
        self._h = h

    def run(self, n_steps: int = 1):
        b = self._batch
        n_st = len(b.n_atoms)
        A = int(b.n_atoms.sum())
        e = np.zeros(n_st)
        f = np.zeros((A, 3))
        s = np.zeros((n_st, 6))
        _check(lib().plm_eval_graph_run(self._h, int(n_steps), _pd(e), _pd(f), _pd(s)))
        off = np.concatenate([[0], np.cumsum(b.n_atoms)])
        return e, [f[off[k]:off[k + 1]] for k in range(n_st)], s

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.plm_eval_graph_destroy(self._h)
            self._h = None


def ridge(xtx, xty, alphas) -> np.ndarray:
    """Coefficients (n_alpha x F) of (xtx + alpha I) c = xty on the GPU (plm_ridge)."""
    A = _d(xtx)
    b = _d(xty)
    al = _d(alphas)
    F = A.shape[0]
    out = np.zeros((len(al), F))
    info = np.zeros(len(al), dtype=np.int32)
    _check(lib().plm_ridge(F, _pd(A), _pd(b), _pd(al), len(al), _pd(out), _pi(info)))
    out[info != 0] = 1e30
    return out


def tuning_report(model: "Model") -> str:
    """The launch shapes the autotuner chose for model so far (plm_model_tuning)."""
    n = ctypes.c_int()
    _check(lib().plm_model_tuning(model.handle, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(lib().plm_model_tuning(model.handle, buf, n.value + 1, None))
    return buf.value.decode()


def feature_attrs(model: "Model"):
    """(radial ids, gtinv ids (-1 for pair features), type-pair combinations,
    polynomial factor lists, type-pair table) of the columns of X
    (FeaturesAttr, py_features_attr.cpp:11-44)."""
    nl, npoly = ctypes.c_int(), ctypes.c_int()
    _check(lib().plm_model_feature_attrs(model.handle, ctypes.byref(nl), ctypes.byref(npoly),
                                         None, None, None, None, None))
    L, P = nl.value, npoly.value
    nt = model.n_type
    radial, gt = np.zeros(max(L, 1), np.int32), np.zeros(max(L, 1), np.int32)
    tc, poly = np.zeros(max(6 * L, 1), np.int32), np.zeros(max(3 * P, 1), np.int32)
    tps = np.zeros(nt * nt, np.int32)
    _check(lib().plm_model_feature_attrs(model.handle, None, None, _pi(radial), _pi(gt), _pi(tc),
                                         _pi(poly), _pi(tps)))
    tcomb = [[int(v) for v in tc[6 * f:6 * f + 6] if v >= 0] for f in range(L)]
    polys = [[int(v) for v in poly[3 * p:3 * p + 3] if v >= 0] for p in range(P)]
    return (radial[:L].tolist(), gt[:L].tolist(), tcomb, polys,
            tps.reshape(nt, nt).tolist())


def solver_cg(xtx, xty, alpha: float, coef0=None, gtol: float = 1e-2, max_iter: int = 10000):
    """(coef, iterations) of the device-resident CG on (xtx + alpha I) c = xty."""
    xtx = _d(xtx)
    xty = _d(xty)
    F = len(xty)
    if xtx.shape != (F, F):
        raise ValueError("xtx must be F x F")
    c0 = None if coef0 is None else _d(coef0)
    if c0 is not None and c0.shape != (F,):
        raise ValueError("coef0 must have F entries")
    out = np.zeros(F)
    n = ctypes.c_int()
    _check(lib().plm_solver_cg(F, _pd(xtx), _pd(xty), ctypes.c_double(alpha), _pd(c0),
                               ctypes.c_double(gtol), int(max_iter), _pd(out), ctypes.byref(n)))
    return out, n.value


def profile_enable(on: bool = True):
    _check(lib().plm_profile_enable(int(on)))


def profile_reset():
    _check(lib().plm_profile_reset())


def profile_get(name: str):
    """(total milliseconds, launches) of a kernel class since the last reset."""
    ms = ctypes.c_double()
    n = ctypes.c_long()
    _check(lib().plm_profile_get(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


# --- surrogate network (FP32 MLP 5-64-64-2, SiLU) ---------------------------

SURR_IN, SURR_HIDDEN, SURR_OUT = 5, 64, 2


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return a.ctypes.data_as(_fp)


def surrogate_n_params() -> int:
    n = ctypes.c_int()
    _check(lib().plm_surrogate_n_params(ctypes.byref(n)))
    return n.value


def surrogate_train(params, m, v, t0: int, x, y, steps: int, lr: float,
                    weight_decay: float = 0.0) -> float:
    """Full-batch Adam(W) steps on the GPU; params, m, v (float32) updated in place."""
    assert params.dtype == np.float32 and m.dtype == np.float32 and v.dtype == np.float32
    x, y = _f(x), _f(y)
    loss = ctypes.c_float()
    _check(lib().plm_surrogate_train(_pf(params), _pf(m), _pf(v), int(t0), _pf(x), _pf(y),
                                     len(x), int(steps), float(lr), float(weight_decay),
                                     ctypes.byref(loss)))
    return loss.value


def surrogate_grad(params, x, y):
    """(gradient of the mean squared loss, loss) at params, on the GPU."""
    params, x, y = _f(params), _f(x), _f(y)
    g = np.zeros_like(params)
    loss = ctypes.c_float()
    _check(lib().plm_surrogate_grad(_pf(params), _pf(x), _pf(y), len(x), _pf(g),
                                    ctypes.byref(loss)))
    return g, loss.value


def surrogate_predict(params, x):
    params, x = _f(params), _f(x)
    y = np.zeros((len(x), SURR_OUT), dtype=np.float32)
    _check(lib().plm_surrogate_predict(_pf(params), _pf(x), len(x), _pf(y)))
    return y


# --- libmlpcpp support functions on the device (get_fn, get_ylm, PolymlpAPI) --
def get_fn(r, params, cutoff: float, precision: int = PLM_FP64):
    """(fn, fn_d), each (len(r), len(params)): radial functions at the
    distances r (libmlpcpp.get_fn, pybind11_mlp.cpp:162-171)."""
    r = _d(np.atleast_1d(r))
    par = _d(params).reshape(-1, 2)
    fn = np.zeros((len(r), len(par)))
    fd = np.zeros_like(fn)
    _check(lib().plm_get_fn(len(r), _pd(r), _pd(par), len(par), float(cutoff), precision,
                            _pd(fn), _pd(fd)))
    return fn, fd


def get_ylm(xyz, lmax: int, precision: int = PLM_FP64):
    """(ylm, dylm) complex: ylm (n, nh), dylm (n, nh, 3) at the points xyz (n, 3)
    (libmlpcpp.get_ylm, pybind11_mlp.cpp:173-184)."""
    xyz = _d(np.atleast_2d(xyz))
    n = len(xyz)
    nh = (lmax + 1) * (lmax + 2) // 2
    y = np.zeros((n, nh, 2))
    dy = np.zeros((n, nh, 3, 2))
    _check(lib().plm_get_ylm(n, _pd(xyz), int(lmax), precision, _pd(y), _pd(dy)))
    return y[..., 0] + 1j * y[..., 1], dy[..., 0] + 1j * dy[..., 1]


def api_features(model: Model, anlm_re, anlm_im, type1: int, precision: int = PLM_FP64):
    """(anlm_full, features) of one atom of type1 from its non-conjugate order
    parameters (PolymlpAPI.compute_anlmtp_conjugate + compute_features)."""
    n_nlm, n_nc, n_lin = model.type_sizes(type1)
    re = _d(anlm_re)[:n_nc]
    im = _d(anlm_im)[:n_nc]
    a = np.zeros((n_nlm, 2))
    f = np.zeros(n_lin)
    _check(lib().plm_api_features(model.handle, int(type1), _pd(re), _pd(im), precision, _pd(a),
                                  _pd(f)))
    return a[:, 0] + 1j * a[:, 1], f


def api_sum_of_prod(model: Model, anlm_re, anlm_im, type1: int, precision: int = PLM_FP64):
    """(prod_sum_e, prod_sum_f) complex per non-conjugate head of one atom of
    type1 (PolymlpAPI.compute_sum_of_prod_anlmtp); the model must hold its
    coefficients."""
    _, n_nc, _ = model.type_sizes(type1)
    re = _d(anlm_re)[:n_nc]
    im = _d(anlm_im)[:n_nc]
    pe = np.zeros((n_nc, 2))
    pf = np.zeros((n_nc, 2))
    _check(lib().plm_api_sum_of_prod(model.handle, int(type1), _pd(re), _pd(im), precision,
                                     _pd(pe), _pd(pf)))
    return pe[:, 0] + 1j * pe[:, 1], pf[:, 0] + 1j * pf[:, 1]
