"""ctypes access to the CPU oracle library oracle/_ref/libplm_ref.so.

Test infrastructure only: tests, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` use it as the checker. The library is the
reference's own C++ hot path compiled from its sources (see oracle/Makefile)
plus ``ref_harness.cpp`` (restatement of model.cpp / py_model.cpp row
assembly).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libplm_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def build(ref_root: str = "/root/reference") -> bool:
    """Compile the oracle library when the reference tree is present."""
    if not os.path.isdir(ref_root):
        return False
    subprocess.run(["make", "-s", "-j8", f"REF_ROOT={ref_root}"], cwd=HERE, check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"oracle library missing: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        L.ref_model_load.restype = ctypes.c_void_p
        L.ref_model_load.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.ref_model_free.argtypes = [ctypes.c_void_p]
        L.ref_model_n_features.argtypes = [ctypes.c_void_p]
        L.ref_model_n_linear.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_neighbor_cell.restype = ctypes.c_long
        L.ref_neighbor_cell.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, _dp,
                                        ctypes.c_long, _dp, _dp]
        L.ref_neighbor.restype = ctypes.c_long
        L.ref_neighbor.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   _ip, _ip, _dp]
        L.ref_structure_rows.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, ctypes.c_int,
                                         ctypes.c_int, _dp, _dp, _dp]
        L.ref_x_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _ip, _ip,
                                  ctypes.c_int, _dp]
        L.ref_eval.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                               _dp, _dp, _dp]
        L.ref_get_fn.restype = None
        L.ref_get_fn.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.ref_get_ylm.restype = None
        L.ref_get_ylm.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.ref_api.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp,
                              _dp, _dp, _ip]
        L.ref_gtinv.restype = ctypes.c_long
        L.ref_gtinv.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                _ip, _ip, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class RefModel:
    """Reference model loaded from a polymlp.yaml file."""

    def __init__(self, yaml_path: str, with_eval: bool = False):
        self._h = lib().ref_model_load(yaml_path.encode(), int(with_eval))
        if not self._h:
            raise RuntimeError(lib().ref_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ref_model_free(self._h)
            self._h = None

    @property
    def n_features(self) -> int:
        return lib().ref_model_n_features(self._h)

    def structure_rows(self, axis, positions_c, types, force=True):
        """(xe[F], xs[6,F], xf[3N,F]) of one structure."""
        axis, pos = _d(axis), _d(positions_c)
        types = np.ascontiguousarray(types, dtype=np.int32)
        n = len(types)
        F = self.n_features
        xe = np.zeros(F)
        xs = np.zeros((6, F))
        xf = np.zeros((3 * n, F))
        rc = lib().ref_structure_rows(self._h, _pd(axis), _pd(pos), _pi(types), n,
                                      int(force), _pd(xe), _pd(xs), _pd(xf))
        if rc:
            raise RuntimeError(lib().ref_last_error().decode())
        return xe, xs, xf

    def x_batch(self, axes, positions_c, types, force=True):
        """X (rows x F) with the PyModel row layout for one dataset."""
        n_st = len(axes)
        n_atoms = np.array([len(t) for t in types], dtype=np.int32)
        ax = _d(np.stack(axes))
        pos = _d(np.concatenate([np.asarray(p).T for p in positions_c]).T.reshape(3, -1))
        # positions are passed per structure as contiguous 3 x N blocks
        pos = _d(np.concatenate([_d(p).ravel() for p in positions_c]))
        tp = np.ascontiguousarray(np.concatenate(types), dtype=np.int32)
        rows = n_st + (6 * n_st + 3 * int(n_atoms.sum()) if force else 0)
        X = np.zeros((rows, self.n_features))
        rc = lib().ref_x_batch(self._h, n_st, _pd(ax), _pd(pos), _pi(tp), _pi(n_atoms),
                               int(force), _pd(X))
        if rc:
            raise RuntimeError("ref_x_batch failed")
        return X

    def eval(self, axis, positions_c, types, use_openmp=True):
        axis, pos = _d(axis), _d(positions_c)
        types = np.ascontiguousarray(types, dtype=np.int32)
        n = len(types)
        e = np.zeros(1)
        f = np.zeros((n, 3))
        s = np.zeros(6)
        rc = lib().ref_eval(self._h, _pd(axis), _pd(pos), _pi(types), n, int(use_openmp),
                            _pd(e), _pd(f), _pd(s))
        if rc:
            raise RuntimeError(lib().ref_last_error().decode())
        return float(e[0]), f, s


def get_fn(r, params, cutoff):
    """(fn, fn_d), each (len(r), len(params)), from the reference's get_fn_."""
    r = _d(np.atleast_1d(r))
    par = _d(params).reshape(-1, 2)
    fn = np.zeros((len(r), len(par)))
    fd = np.zeros_like(fn)
    lib().ref_get_fn(len(r), _pd(r), _pd(par), len(par), float(cutoff), _pd(fn), _pd(fd))
    return fn, fd


def get_ylm(xyz, lmax):
    """(ylm (n, nh), dylm (n, nh, 3)) complex from the reference's get_ylm_."""
    xyz = _d(np.atleast_2d(xyz))
    n = len(xyz)
    nh = (lmax + 1) * (lmax + 2) // 2
    y = np.zeros((n, nh, 2))
    dy = np.zeros((n, nh, 3, 2))
    lib().ref_get_ylm(n, _pd(xyz), int(lmax), _pd(y), _pd(dy))
    return y[..., 0] + 1j * y[..., 1], dy[..., 0] + 1j * dy[..., 1]


def api(model: "RefModel", anlm_re, anlm_im, type1, with_prod=False):
    """PolymlpAPI on one atom: (anlm_full, features[, prod_sum_e, prod_sum_f])
    from its non-conjugate order parameters."""
    re, im = _d(anlm_re), _d(anlm_im)
    n = np.zeros(3, dtype=np.int32)
    L = lib()
    if L.ref_api(model._h, int(type1), _pd(re), _pd(im), len(re), None, None, None, None, _pi(n)):
        raise RuntimeError(L.ref_last_error().decode())
    a = np.zeros((n[0], 2))
    f = np.zeros(n[1])
    pe = pf = None
    if with_prod:
        pe = np.zeros((n[0], 2))
        pf = np.zeros((n[0], 2))
    if L.ref_api(model._h, int(type1), _pd(re), _pd(im), len(re), _pd(a), _pd(f),
                 _pd(pe) if with_prod else None, _pd(pf) if with_prod else None, _pi(n)):
        raise RuntimeError(L.ref_last_error().decode())
    out = (a[:, 0] + 1j * a[:, 1], f)
    if with_prod:
        k = n[2]
        out += (pe[:k, 0] + 1j * pe[:k, 1], pf[:k, 0] + 1j * pf[:k, 1])
    return out


def neighbor_cell(axis, positions_c, cutoff):
    """(translations[n_T,3], positions_c_rev[3,N], axis_rev[3,3]) of NeighborCell."""
    axis, pos = _d(axis), _d(positions_c)
    n = pos.shape[1]
    nt = lib().ref_neighbor_cell(_pd(axis), _pd(pos), n, cutoff, None, 0, None, None)
    tr = np.zeros((nt, 3))
    prev = np.zeros((3, n))
    axr = np.zeros((3, 3))
    lib().ref_neighbor_cell(_pd(axis), _pd(pos), n, cutoff, _pd(tr), nt, _pd(prev), _pd(axr))
    return tr, prev, axr


def neighbor_list(axis, positions_c, cutoff, half=False):
    """CSR neighbour list (offset[N+1], neigh[M], d[M,3]) with d = r_j - r_i + T."""
    axis, pos = _d(axis), _d(positions_c)
    n = pos.shape[1]
    m = lib().ref_neighbor(_pd(axis), _pd(pos), n, cutoff, int(half), None, None, None)
    off = np.zeros(n + 1, dtype=np.int32)
    nb = np.zeros(m, dtype=np.int32)
    d = np.zeros((m, 3))
    lib().ref_neighbor(_pd(axis), _pd(pos), n, cutoff, int(half), _pi(off), _pi(nb), _pd(d))
    return off, nb, d


def gtinv(order, maxl, version=1):
    """(l_comb, lm_seq, lm_coeffs) from the reference's Readgtinv."""
    ml = np.ascontiguousarray(maxl, dtype=np.int32)
    L = lib()
    n = L.ref_gtinv(order, _pi(ml), len(ml), version, None, None, None, None, None)
    orders = np.zeros(n, dtype=np.int32)
    nt = np.zeros(n, dtype=np.int32)
    L.ref_gtinv(order, _pi(ml), len(ml), version, _pi(orders), None, _pi(nt), None, None)
    l_flat = np.zeros(int(orders.sum()), dtype=np.int32)
    lm_flat = np.zeros(int((orders * nt).sum()), dtype=np.int32)
    c_flat = np.zeros(int(nt.sum()))
    L.ref_gtinv(order, _pi(ml), len(ml), version, _pi(orders), _pi(l_flat), _pi(nt),
                _pi(lm_flat), _pd(c_flat))
    l_comb, lm_seq, coeffs = [], [], []
    pl = pm = pc = 0
    for o, k in zip(orders, nt):
        l_comb.append(l_flat[pl:pl + o].tolist())
        pl += o
        lm_seq.append(lm_flat[pm:pm + o * k].reshape(k, o).tolist())
        pm += o * k
        coeffs.append(c_flat[pc:pc + k].tolist())
        pc += k
    return l_comb, lm_seq, coeffs
