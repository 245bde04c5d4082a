"""GPU ridge solve and the whole fit, against the oracle.

plm_ridge (blocked FP64 Cholesky on FP64 MFMA plus one refinement step)
against LAPACK posv as solvers.py:9-84 calls it:
  - backward error ||(A + a I) c - b|| / (||A + a I|| ||c||) <= 1e-14;
  - forward difference to posv <= 100 * cond * eps (relative, normwise),
    where cond is the 2-norm condition number of A + a I;
  - a matrix that is not positive definite gives 1e30 coefficients like the
    reference's failed posv.

Whole fit on cfg1 (Morse targets): FP64 GPU fit against the oracle's fit
(reference X from the compiled sources, numpy X^T X, the same scales, posv):
predicted test energies within 1e-6 eV/atom and forces within 1e-5 eV/A.
Mixed precision against FP64: within 0.1 meV/atom and 1 meV/A.
"""

import sys

import numpy as np
import pytest

from oracle import ref
from oracle.numerics import solver_ridge_posv
from pypolymlp_amd import capi, datasets, mlp_dev

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spd(F, rows, seed, col_scale_decades=3.0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(rows, F)) * np.logspace(0, -col_scale_decades, F)[None, :]
    X[:, 1:] += 0.3 * X[:, :-1]
    A = X.T @ X
    b = X.T @ rng.normal(size=rows)
    return A, b


@pytest.mark.parametrize("F", [61, 300, 1000])
def test_ridge_vs_posv(F):
    A, b = _spd(F, 3 * F, seed=F)
    alphas = np.array([1e-6, 1e-4, 1e-2]) * np.trace(A) / F
    say(f"ridge F={F}")
    C = mlp_dev.solver_ridge(A, b, alphas)
    Cr = solver_ridge_posv(A, b, alphas)
    eps = np.finfo(np.float64).eps
    for k, a in enumerate(alphas):
        Aa = A + a * np.eye(F)
        c = C[:, k]
        bwd = np.linalg.norm(Aa @ c - b) / (np.linalg.norm(Aa, 2) * np.linalg.norm(c))
        cond = np.linalg.cond(Aa)
        fwd = np.linalg.norm(c - Cr[:, k]) / np.linalg.norm(Cr[:, k])
        say(f"  alpha {a:.2e}: backward {bwd:.2e}, cond {cond:.1e}, fwd vs posv {fwd:.2e}")
        assert bwd <= 1e-14
        assert fwd <= max(100 * cond * eps, 1e-13)


def test_ridge_not_positive_definite():
    A, b = _spd(130, 400, seed=3)
    A[70, 70] = -abs(A[70, 70])
    say("ridge non-PD")
    C = mlp_dev.solver_ridge(A, b, np.array([1e-8]))
    assert np.all(C == 1e30)
    Cr = solver_ridge_posv(A, b, np.array([1e-8]))
    assert np.all(Cr == 1e30)


def _oracle_normal(R, ds, min_e, scales=None, weight_stress=0.1):
    """Oracle normal equations: reference X, numpy X^T X, the reference's
    scale rule (data_sequential.py:72-92)."""
    ds.sort_by_atoms()
    sts = ds.structures
    X = R.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                  [s.types for s in sts], force=True)
    w, y = mlp_dev.row_weights(ds, min_e, weight_stress)
    ne = len(sts)
    xe_sum = X[:ne].sum(axis=0)
    xe_sq = (X[:ne] ** 2).sum(axis=0)
    scales, zero = mlp_dev.compute_scales(scales, xe_sum, xe_sq, ne)
    Xw = X * w[:, None]
    xtx = Xw.T @ Xw
    xty = Xw.T @ y
    xtx[zero] = 0.0
    xtx[:, zero] = 0.0
    xty[zero] = 0.0
    xtx /= scales[:, None] * scales[None, :]
    xty /= scales
    return mlp_dev.NormalEquations(xtx=xtx, xty=xty, y_sq_norm=float(y @ y),
                                   total_n_data=X.shape[0], scales=scales, min_energy=min_e)


def _oracle_fit(R, train, test, alphas, min_e, weight_stress=0.1):
    """Oracle fit: _oracle_normal, posv, model selection on the test RMSE."""
    def normal(ds, scales=None):
        return _oracle_normal(R, ds, min_e, scales, weight_stress)

    tr = normal(train)
    coefs = solver_ridge_posv(tr.xtx, tr.xty, alphas)
    te = normal(test, scales=tr.scales.copy())
    rm = mlp_dev.compute_rmse(coefs, te)
    k = int(np.argmin(rm))
    return coefs[:, k] / tr.scales, float(alphas[k])


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
def test_fit_cfg1_fp64_vs_oracle_and_mixed(tmp_path):
    from tests.helpers import write_model_yaml

    spec = datasets.config1_model()
    full = datasets.config1_dataset(48, seed=11)
    train = datasets.Dataset(structures=full.structures[:36], energies=full.energies[:36],
                             forces=full.forces[:36], stresses=full.stresses[:36], name="train")
    test = datasets.Dataset(structures=full.structures[36:], energies=full.energies[36:],
                            forces=full.forces[36:], stresses=full.stresses[36:], name="test")
    alphas = np.array([1e-5, 1e-4, 1e-3, 1e-2])
    m = capi.Model(spec.as_dict())
    say("fit cfg1 FP64")
    r64 = mlp_dev.fit(m, [train], [test], alphas, precision=capi.PLM_FP64)
    say(f"  alpha {r64.alpha}, rmse train {r64.rmse_train:.4e}, test {r64.rmse_test:.4e}")
    say("fit cfg1 MIXED")
    rmx = mlp_dev.fit(m, [train], [test], alphas, precision=capi.PLM_MIXED)
    say(f"  alpha {rmx.alpha}, rmse train {rmx.rmse_train:.4e}, test {rmx.rmse_test:.4e}")
    path = write_model_yaml(spec, str(tmp_path / "m.yaml"))
    R = ref.RefModel(path)
    say("fit cfg1 oracle")
    c_ref, a_ref = _oracle_fit(R, train, test, alphas, r64_min_e(train))
    assert a_ref == r64.alpha

    def predict(coeffs):
        m.set_coeffs(coeffs)
        sts = test.structures
        b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                       [s.types for s in sts])
        return capi.eval_batch(m, b, capi.PLM_FP64)

    n_at = np.array([s.n_total for s in test.structures])
    E64, F64, _ = predict(r64.coeffs_unscaled)
    Er, Fr, _ = predict(c_ref)
    Em, Fm, _ = predict(rmx.coeffs_unscaled)
    dref_e = np.max(np.abs(E64 - Er) / n_at)
    dref_f = max(np.abs(a - b).max() for a, b in zip(F64, Fr))
    dmix_e = np.max(np.abs(Em - E64) / n_at)
    dmix_f = max(np.abs(a - b).max() for a, b in zip(Fm, F64))
    say(f"  FP64 vs oracle fit: dE {dref_e * 1e3:.3e} meV/atom, dF {dref_f * 1e3:.3e} meV/A")
    say(f"  mixed vs FP64 fit: dE {dmix_e * 1e3:.3e} meV/atom, dF {dmix_f * 1e3:.3e} meV/A")
    assert dref_e < 1e-6
    assert dref_f < 1e-5
    assert dmix_e < 1e-4
    assert dmix_f < 1e-3


def r64_min_e(ds):
    return mlp_dev.get_min_energy([ds])


def test_fit_cfg2_mixed_predictions():
    """Binary-alloy model (F = 13,020) fitted on 108 structures of the seeded
    binary fcc alloy set: the mixed fit predicts the 12 test structures within
    0.1 meV/atom and 1 meV/A of the FP64 fit, at each alpha."""
    ds = datasets.config2_dataset(120)
    train, test = ds.split(0.9)
    m = capi.Model(datasets.config2_model().as_dict())
    alphas = np.array([1e-3, 1e-1])
    fits = {}
    for prec in (capi.PLM_FP64, capi.PLM_MIXED):
        say(f"cfg2 fit prec={prec}")
        ne = mlp_dev.calc_xtx_xty(m, [train], precision=prec)
        fits[prec] = mlp_dev.solver_ridge(ne.xtx, ne.xty, alphas) / ne.scales[:, None]
    sts = test.structures
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    n = np.array([s.n_total for s in sts])
    for k, a in enumerate(alphas):
        pred = {}
        for prec, c in fits.items():
            m.set_coeffs(c[:, k])
            E, F, _ = capi.eval_batch(m, b, capi.PLM_FP64)
            pred[prec] = (E / n, np.concatenate([f.reshape(-1) for f in F]))
        de = np.abs(pred[capi.PLM_MIXED][0] - pred[capi.PLM_FP64][0]).max()
        df = np.abs(pred[capi.PLM_MIXED][1] - pred[capi.PLM_FP64][1]).max()
        say(f"  alpha {a:g}: max |dE| {de * 1e3:.3e} meV/atom, max |dF| {df * 1e3:.3e} meV/A")
        assert de <= 1e-4 and df <= 1e-3


def test_fit_device_resident_matches_host_path():
    """The device-resident solve (X^T X scaled and solved on the GPU,
    plm_normal_ridge / plm_normal_sq_error) against the reference-API path on
    the same accumulator (its X^T X read back and scaled on the host as
    calc_xtx_xty does, then solver_ridge and compute_rmse): the scaled
    systems are bit-identical, so the coefficients agree to 1e-12 and the
    RMSEs to 1e-10 (relative). Dropped columns (a zero mask) against numpy on
    the read-back normal equations. fit() picks the same alpha as the host
    path (the energy-row sums use device atomics, so scales from two
    accumulations differ in the last bits; coefficients then agree to the
    solve's conditioning)."""
    say("device-resident fit vs host path (cfg1 model, 30 + 10 structures)")
    spec = datasets.config1_model()
    full = datasets.config1_dataset(40, seed=13)
    train, test = full.subset(range(30), "train"), full.subset(range(30, 40), "test")
    model = capi.Model(spec.as_dict())
    alphas = np.array([1e-4, 1e-2, 1.0])
    d = mlp_dev.accumulate_normal(model, [train])
    tr = mlp_dev._finish_normal(d.acc.result(), len(train.energies), None, True, 1e-10,
                                d.min_energy, n_rows=d.total_n_data)
    assert np.array_equal(d.scales, tr.scales)
    coefs = mlp_dev.solver_ridge(tr.xtx, tr.xty, alphas)
    rmse_tr = mlp_dev.compute_rmse(coefs, tr)
    c2, info, sq = d.acc.ridge(d.scales, d.zero, alphas)
    assert np.all(info == 0)
    for k in range(len(alphas)):
        assert np.linalg.norm(c2[:, k] - coefs[:, k]) <= 1e-12 * np.linalg.norm(coefs[:, k])
    # the squared-error sum c^T A c - 2 c^T b + y^T y cancels: its rounding is
    # bounded by the size of its terms, in any summation order
    def sq_close(sq_dev, ne, C):
        for k in range(C.shape[1]):
            c = C[:, k]
            cac, cb = c @ (ne.xtx @ c), c @ ne.xty
            host = cac - 2 * cb + ne.y_sq_norm
            assert abs(sq_dev[k] - host) <= 1e-12 * (abs(cac) + 2 * abs(cb) + ne.y_sq_norm)

    sq_close(sq, tr, coefs)
    np.testing.assert_allclose(mlp_dev._rmse_from_sq(sq, d.total_n_data), rmse_tr, rtol=1e-5)
    dt = mlp_dev.accumulate_normal(model, [test], scales=d.scales.copy(), min_energy=d.min_energy)
    te = mlp_dev._finish_normal(dt.acc.result(), len(test.energies), d.scales.copy(), True, 1e-10,
                                d.min_energy, n_rows=dt.total_n_data)
    rmse_te = mlp_dev.compute_rmse(coefs, te)
    sq_te = dt.acc.sq_error(dt.scales, dt.zero, c2)
    sq_close(sq_te, te, coefs)
    np.testing.assert_allclose(mlp_dev._rmse_from_sq(sq_te, dt.total_n_data), rmse_te, rtol=1e-5)
    k = int(np.argmin(rmse_te))
    fr = mlp_dev.fit(model, [train], [test], alphas)
    assert fr.alpha == alphas[k]
    Aa = tr.xtx + alphas[k] * np.eye(len(tr.xty))
    bound = max(100 * np.linalg.cond(Aa) * np.finfo(float).eps, 1e-12)
    assert np.linalg.norm(fr.coeffs - coefs[:, k]) <= bound * np.linalg.norm(coefs[:, k])
    # dropped columns: the device zeroes rows and columns of the mask
    zero = np.zeros(len(d.scales), dtype=bool)
    zero[[3, 100, 2000]] = True
    a = 1.0
    c3, info, _ = d.acc.ridge(d.scales, zero, [a])
    xtx, xty = tr.xtx.copy(), tr.xty.copy()
    xtx[zero] = 0.0
    xtx[:, zero] = 0.0
    xty[zero] = 0.0
    Aa = xtx + a * np.eye(len(xty))
    c_ref = np.linalg.solve(Aa, xty)
    assert np.all(c3[zero, 0] == 0.0)
    bound = max(100 * np.linalg.cond(Aa) * np.finfo(float).eps, 1e-12)
    assert np.linalg.norm(c3[:, 0] - c_ref) <= bound * np.linalg.norm(c_ref)


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
def test_fp64_coefficients_cfg1_vs_oracle(tmp_path):
    """The FP64 fit's coefficients on cfg1's own normal matrix (F = 2277): the
    device path (plm_normal_accumulate, plm_normal_ridge) against the
    reference X with numpy normal equations and LAPACK, normwise within
    max(1e-10, 100 cond eps), and within 1e-10 wherever cond eps < 1e-12."""
    from tests.helpers import write_model_yaml

    spec = datasets.config1_model()
    ds = datasets.config1_dataset(24, seed=4)
    m = capi.Model(spec.as_dict())
    min_e = mlp_dev.get_min_energy([ds])
    alphas = np.array([1e-3, 1e-1, 10.0])
    say("cfg1 FP64 coefficients: 24 structures, device normal equations and ridge")
    dn = mlp_dev.accumulate_normal(m, [ds], min_energy=min_e)
    c_dev, info, _ = dn.acc.ridge(dn.scales, dn.zero, alphas)
    assert np.all(info == 0)
    R = ref.RefModel(write_model_yaml(spec, str(tmp_path / "m.yaml")))
    ne = _oracle_normal(R, ds, min_e)
    np.testing.assert_allclose(dn.scales, ne.scales, rtol=1e-10)
    for k, a in enumerate(alphas):
        A = ne.xtx + a * np.eye(len(ne.xty))
        c_ref = np.linalg.solve(A, ne.xty)
        cond = np.linalg.cond(A)
        err = np.linalg.norm(c_dev[:, k] - c_ref) / np.linalg.norm(c_ref)
        say(f"  alpha {a:g}: cond {cond:.2e}, relative coefficient error {err:.2e}")
        assert err <= max(1e-10, 100 * cond * np.finfo(float).eps)
        if cond * np.finfo(float).eps < 1e-12:
            assert err <= 1e-10
