"""Radial basis, spherical harmonics and the per-atom invariant contractions
pinned directly by the reference's own known answers.

Golden values: tests/test_cxx/test_functions.py:14-111 (get_fn at r = 1.2,
get_ylm tables at one point) and tests/test_cxx/test_polymlp_api.py:52-105
(synthetic order parameters anlm = arange(n) (1 + i) for the MgO gtinv model
polymlp.yaml.MgO: conjugate fill, feature values, prod_sum_e / prod_sum_f).

CPU: the oracle (the reference's get_fn_ / get_ylm_ / PolymlpAPI compiled
from its sources) against those values, plus closed forms: the radial
derivative against central differences, Y_lm against scipy's sph_harm with
the reference's phase and the derivatives against central differences.
GPU: the device functions of the X and evaluation kernels (plm_get_fn,
plm_get_ylm, plm_api_features, plm_api_sum_of_prod) against the same values
and against the oracle at random points: FP64 to 1e-12 (relative to the
largest entry), FP32 (the mixed path) to 1e-5.

The reference's synthetic compute_features_deriv test (derivatives given for
all 450 components, conjugate ones included, with values that do not satisfy
the conjugate relation) has no device counterpart: the device folds each
conjugate component's derivative into its non-conjugate partner, which is
exact only for consistent derivatives. The derivative rows are pinned by the
X parity tests and the finite-difference tests of the oracle instead.
"""

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi
from tests.helpers import load_model, ref_file

oracle = pytest.mark.skipif(not ref.available(), reason="oracle library not built")

FN_PARAMS = [[1.0, 0.0], [1.0, 1.0], [1.0, 2.0]]
FN_TRUE = [0.21430317094756238, 0.8690422117212636, 0.4769404780495180]
FN_D_TRUE = [-0.5507864848009192, -0.495464911460655, 0.6819640474131319]
XYZ = (0.173723561607389, 0.446843340790007, 0.877582561890373)
YLM_SUMS_L10 = (-3.094632553138235 + 0.09510814961092404j, -6.916830463136405 - 1.910490992920798j,
                12.686387327323297 + 23.645024627283863j, -5.090360117438893 - 11.661266919165213j)
YLM_L3 = [
    0.2820947917738782 + 0j, 0.060020474160676 - 0.15438175997317355j, 0.4287890441418357 + 0j,
    -0.0654692650582695 - 0.059970782327791j, 0.11778023300234376 - 0.3029486173713942j,
    0.41330596756220733 + 0j, -0.04122949872902822 + 0.02034538591673029j,
    -0.15201080910457251 - 0.13924407332460548j, 0.1600528676431333 - 0.4116802431342542j,
    0.2786165933635158 + 0j]
YLM_DX_L3 = [
    0j, 0.33506717893077853 + 0.026819749189756756j, -0.07449075992654758 + 0j,
    0.15695696806852266 - 0.3243714339893415j, 0.6370526044791094 + 0.10525862558758538j,
    -0.2531838614701831 + 0j, -0.19065705919150708 - 0.20493096358473054j,
    0.3908407833648457 - 0.7289567725435584j, 0.8183858345809413 + 0.26473083057567426j,
    -0.48656634670427024 + 0j]
YLM_DY_L3 = [
    0j, -0.026819749189756728 - 0.27650968808788157j, -0.1916015289784915 + 0j,
    -0.2866990995506649 - 0.08061477081476034j, -0.10525862558758535 - 0.4072338630594455j,
    -0.6512272799767366 + 0j, -0.139058164020008 + 0.18487106464090822j,
    -0.5977516482661335 - 0.12495634515471779j, -0.26473083057567415 - 0.24037993759857124j,
    -1.2515224179474653 + 0j]
YLM_DZ_L3 = [
    0j, -0.05267292147980093 + 0.13548274042640215j, 0.11230472403440366 + 0j,
    0.11490937070983197 + 0.10525862558758538j, -0.07251389697144246 + 0.18651673772179642j,
    0.38170833161604206 + 0j, 0.10854686736022928 - 0.053564267686357386j,
    0.2269907120047056 + 0.20792673581945217j, -0.027210765396762822 + 0.06999021435458216j,
    0.733563455990153 + 0j]
# PolymlpAPI on polymlp.yaml.MgO with anlm = arange (1 + i)
API_CONJ_SUM = 31581 + 17199j
API_FEATURE_SUM = 849391612.5622855
API_PROD_E_SUM = -24266095.550829127 - 955759.7528607883j
API_PROD_F_SUM = -72763009.4963038 - 2902187.696305793j


def _random_points(n, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v *= rng.uniform(0.5, 7.0, size=(n, 1)) / np.linalg.norm(v, axis=1, keepdims=True)
    v[0] = [0.0, 0.0, 1.3]     # on the z axis: rho = 0 branch
    v[1] = [0.0, 0.0, -2.1]
    v[2] = [1.1, 0.0, 0.0]     # theta = pi / 2
    return v


def _ylm_scipy(xyz, lmax):
    """Y_l^{-m} (m >= 0) at key l(l+3)/2 - m from scipy with the reference's
    phase: Y_l^{-m} = (-1)^m conj(Y_l^m) of the Condon-Shortley convention."""
    from scipy.special import sph_harm_y
    x, y, z = xyz
    r = np.sqrt(x * x + y * y + z * z)
    theta, phi = np.arccos(z / r), np.arctan2(y, x)
    out = np.zeros((lmax + 1) * (lmax + 2) // 2, dtype=complex)
    for l in range(lmax + 1):
        for m in range(l + 1):
            out[l * (l + 3) // 2 - m] = (-1) ** m * np.conj(sph_harm_y(l, m, theta, phi))
    return out


# --- the oracle against the reference's known answers (CPU) ------------------
@oracle
def test_oracle_get_fn_golden_and_derivative():
    fn, fd = ref.get_fn([1.2], FN_PARAMS, 6.0)
    np.testing.assert_allclose(fn[0], FN_TRUE, rtol=1e-12)
    np.testing.assert_allclose(fd[0], FN_D_TRUE, rtol=1e-12)
    r = np.linspace(0.3, 5.9, 40)
    h = 1e-6
    fp, _ = ref.get_fn(r + h, FN_PARAMS, 6.0)
    fm, _ = ref.get_fn(r - h, FN_PARAMS, 6.0)
    _, d = ref.get_fn(r, FN_PARAMS, 6.0)
    np.testing.assert_allclose((fp - fm) / (2 * h), d, atol=1e-8)


@oracle
def test_oracle_get_ylm_golden_scipy_and_derivatives():
    y, dy = ref.get_ylm([XYZ], 3)
    np.testing.assert_allclose(y[0], YLM_L3, atol=1e-7)
    np.testing.assert_allclose(dy[0, :, 0], YLM_DX_L3, atol=1e-7)
    np.testing.assert_allclose(dy[0, :, 1], YLM_DY_L3, atol=1e-7)
    np.testing.assert_allclose(dy[0, :, 2], YLM_DZ_L3, atol=1e-7)
    y, dy = ref.get_ylm([XYZ], 10)
    assert y.shape == (1, 66)
    for k, s in enumerate(YLM_SUMS_L10):
        got = y[0].sum() if k == 0 else dy[0, :, k - 1].sum()
        assert got == pytest.approx(s, rel=1e-6)
    pts = _random_points(12, seed=1)[3:]
    y, dy = ref.get_ylm(pts, 8)
    for k, p in enumerate(pts):
        np.testing.assert_allclose(y[k], _ylm_scipy(p, 8), atol=1e-12)
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        fd = (ref.get_ylm(pts + e, 8)[0] - ref.get_ylm(pts - e, 8)[0]) / (2 * h)
        np.testing.assert_allclose(fd, dy[:, :, a], atol=1e-7)


@pytest.fixture(scope="module")
def mgo_api():
    spec, coeffs = load_model("polymlp.yaml.MgO")
    return spec, coeffs


@oracle
def test_oracle_polymlp_api_golden():
    R = ref.RefModel(ref_file("polymlp.yaml.MgO"), with_eval=True)
    n = 1899
    a0, f0, pe, pf = ref.api(R, np.arange(n), np.arange(n), 0, with_prod=True)
    a1, f1 = ref.api(R, np.arange(n), np.arange(n), 1)
    assert a0.shape == a1.shape == (450,)
    assert a0.sum() == API_CONJ_SUM and a1.sum() == API_CONJ_SUM
    assert f0.shape == f1.shape == (891,)
    assert f0.sum() == pytest.approx(API_FEATURE_SUM)
    assert f1.sum() == pytest.approx(API_FEATURE_SUM)
    assert pe.shape == pf.shape == (270,)
    assert pe.sum() == pytest.approx(API_PROD_E_SUM)
    assert pf.sum() == pytest.approx(API_PROD_F_SUM)


# --- the device functions (GPU) ----------------------------------------------
def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(capi.PLM_FP64, 1e-12), (capi.PLM_MIXED, 1e-5)])
def test_device_get_fn(precision, tol):
    print(f"get_fn precision {precision}", flush=True)
    fn, fd = capi.get_fn([1.2], FN_PARAMS, 6.0, precision)
    assert _close(fn[0], FN_TRUE, tol) and _close(fd[0], FN_D_TRUE, tol)
    r = np.random.default_rng(2).uniform(0.05, 7.99, 300)
    par = [[b, m] for b in (0.5, 1.0, 2.0) for m in np.linspace(0.0, 7.0, 8)] + [[0.0, 0.0]]
    fn, fd = capi.get_fn(r, par, 8.0, precision)
    rfn, rfd = ref.get_fn(r, par, 8.0) if ref.available() else (None, None)
    if rfn is not None:
        assert _close(fn, rfn, tol) and _close(fd, rfd, 10 * tol)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(capi.PLM_FP64, 1e-12), (capi.PLM_MIXED, 1e-5)])
def test_device_get_ylm(precision, tol):
    print(f"get_ylm precision {precision}", flush=True)
    y, dy = capi.get_ylm([XYZ], 3, precision)
    assert _close(y[0], YLM_L3, max(tol, 1e-9))
    assert _close(dy[0, :, 0], YLM_DX_L3, max(tol, 1e-9))
    assert _close(dy[0, :, 1], YLM_DY_L3, max(tol, 1e-9))
    assert _close(dy[0, :, 2], YLM_DZ_L3, max(tol, 1e-9))
    y, dy = capi.get_ylm([XYZ], 10, precision)
    for k, s in enumerate(YLM_SUMS_L10):
        got = y[0].sum() if k == 0 else dy[0, :, k - 1].sum()
        assert got == pytest.approx(s, rel=max(tol, 1e-9))
    if ref.available():
        pts = _random_points(200, seed=3)
        for lmax in (4, 8, 12):
            y, dy = capi.get_ylm(pts, lmax, precision)
            ry, rdy = ref.get_ylm(pts, lmax)
            assert _close(y, ry, tol)
            # derivatives scale as 1/r: compare per point
            for k in range(len(pts)):
                assert _close(dy[k], rdy[k], 10 * tol), (lmax, k)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(capi.PLM_FP64, 1e-12), (capi.PLM_MIXED, 1e-5)])
def test_device_polymlp_api(mgo_api, precision, tol):
    """Conjugate fill, feature values (the W' / value kernel) and prod_sum (the
    evaluation's per-atom kernel) of synthetic order parameters."""
    print(f"polymlp api precision {precision}", flush=True)
    spec, coeffs = mgo_api
    model = capi.Model(spec.as_dict())
    model.set_coeffs(coeffs)
    n = 1899
    for t in (0, 1):
        a, f = capi.api_features(model, np.arange(n), np.arange(n), t, precision)
        assert a.shape == (450,) and f.shape == (891,)
        assert a.sum() == API_CONJ_SUM
        assert f.sum() == pytest.approx(API_FEATURE_SUM, rel=max(tol, 1e-12))
    pe, pf = capi.api_sum_of_prod(model, np.arange(n), np.arange(n), 0, precision)
    assert pe.shape == pf.shape == (270,)
    assert pe.sum() == pytest.approx(API_PROD_E_SUM, rel=max(tol, 1e-12))
    assert pf.sum() == pytest.approx(API_PROD_F_SUM, rel=max(tol, 1e-12))
    if ref.available():
        # elementwise against the oracle on random order parameters of both types
        R = ref.RefModel(ref_file("polymlp.yaml.MgO"), with_eval=True)
        rng = np.random.default_rng(4)
        for t in (0, 1):
            re, im = rng.normal(size=270), rng.normal(size=270)
            a, f = capi.api_features(model, re, im, t, precision)
            ra, rf, rpe, rpf = ref.api(R, re, im, t, with_prod=True)
            assert _close(a, ra, 1e-7 if precision else 0.0)
            assert _close(f, rf, tol)
            pe, pf = capi.api_sum_of_prod(model, re, im, t, precision)
            assert _close(pe, rpe, 10 * tol) and _close(pf, rpf, 10 * tol)
