"""Golden values of the reference's Si mp-149 phono3py dataset on the GPU path.

Dataset: tests/files/phonopy_training_dataset.yaml.xz with
polymlp.in.phono3py.Si (200 snapshots of a 64-atom supercell, 180 train /
20 test), model polymlp.yaml.Si for evaluation. Expected numbers and
tolerances are the reference's own:
  X sums          tests/test_mlp_dev/test_core_features.py:13-31 (rel 1e-6)
  calc_xy         tests/test_mlp_dev/test_core_data.py:13-36
  calc_xtx_xty    tests/test_mlp_dev/test_core_data.py:39-71 (rel 1e-6)
  ridge           tests/test_mlp_dev/test_standard_solvers.py:13-37 (rtol 1e-3)
  eval accuracy   tests/test_mlp_dev/test_core_eval_accuracy.py:12-19
"""

import sys

import numpy as np
import pytest

from pypolymlp_amd import capi, mlp_dev
from pypolymlp_amd.dataset_io import load_datasets
from pypolymlp_amd.params import parse_polymlp_in
from tests.helpers import FILES, load_model, ref_file

pytestmark = [pytest.mark.gpu]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def si():
    inp = parse_polymlp_in(ref_file("polymlp.in.phono3py.Si"))
    train, test = load_datasets(inp, prefix=FILES)
    return inp, capi.Model(inp.spec.as_dict()), train, test


@pytest.fixture
def xrows_path(request, monkeypatch):
    if request.param == "generic":
        monkeypatch.setenv("PLM_XROWS", "generic")
    else:
        monkeypatch.delenv("PLM_XROWS", raising=False)
    return request.param


@pytest.mark.parametrize("xrows_path", ["group", "generic"], indirect=True)
def test_si_features_x_sums(si, xrows_path):
    inp, m, train, _ = si
    say(f"si features_x path={xrows_path}")
    feat = mlp_dev.compute_features(m, train, capi.PLM_FP64)
    x = feat.x
    assert x.shape == (35820, 168)
    assert feat.n_data == (180, 34560, 180 * 6)
    assert feat.first_indices[0] == (0, 1260, 180)
    assert np.sum(x) == pytest.approx(5165294.450079148, rel=1e-6)
    assert np.sum(x[:, :20]) == pytest.approx(447.7438322711305, rel=1e-6)
    assert np.sum(x[:, 20:40]) == pytest.approx(15803.28147846774, rel=1e-6)
    assert np.sum(x[:, 40:60]) == pytest.approx(93568.30539681061, rel=1e-6)
    assert np.sum(x[:, 60:80]) == pytest.approx(308321.82717270905, rel=1e-6)
    assert np.sum(x[:, -60:-40]) == pytest.approx(1987480.894857876, rel=1e-6)
    assert np.sum(x[:, -40:-20]) == pytest.approx(34372.80650408206, rel=1e-6)
    assert np.sum(x[:, -20:]) == pytest.approx(2008586.8132866116, rel=1e-6)


def test_si_features_x_mixed_close_to_fp64(si):
    """Mixed X (FP32 stream) of the same rows: per column, the L1 norm of the
    difference within 1e-4 of the column's L1 norm (FP32 rounding amplified by
    cancellation in the force rows of high-order invariants); the fit-level
    tolerance of the mixed path is checked by the ridge coefficients below."""
    _, m, train, _ = si
    say("si features_x mixed")
    x64 = mlp_dev.compute_features(m, train, capi.PLM_FP64).x
    x32 = mlp_dev.compute_features(m, train, capi.PLM_MIXED).x
    ratio = np.abs(x32 - x64).sum(axis=0) / np.abs(x64).sum(axis=0)
    say(f"  worst column {int(np.argmax(ratio))}: relative L1 {ratio.max():.2e}, "
        f"median {np.median(ratio):.2e}")
    assert ratio.max() <= 1e-4


def test_si_calc_xy(si):
    _, m, train, _ = si
    say("si calc_xy")
    d = mlp_dev.calc_xy(m, train, scale_threshold=1e-10)
    assert d.x.shape == (35820, 168)
    assert d.y.shape[0] == 35820 and d.weights.shape[0] == 35820
    assert d.scales.shape[0] == 168
    i, f = 34500, 56
    assert d.x[i, f] == pytest.approx(-1.6856145695486902, abs=1e-8)
    assert d.y[i] == pytest.approx(0.11092374)
    assert d.weights[i] == pytest.approx(1.0)
    assert d.scales[f] == pytest.approx(0.0032488632685359524)
    assert d.min_energy == pytest.approx(-5.737324395625)
    assert d.n_structures == 180
    assert d.first_indices[0] == (0, 1260, 180)


@pytest.mark.parametrize("batch_size", [4096, 10])
def test_si_calc_xtx_xty(si, batch_size):
    _, m, train, _ = si
    say(f"si calc_xtx_xty batch={batch_size}")
    ne = mlp_dev.calc_xtx_xty(m, train, precision=capi.PLM_FP64, batch_size=batch_size)
    assert ne.xtx.shape == (168, 168) and ne.xty.shape[0] == 168
    assert ne.xty[56] == pytest.approx(6.899130774433e5, rel=1e-6)
    assert ne.scales[56] == pytest.approx(0.0032488632685359524)
    assert ne.min_energy == pytest.approx(-5.737324395625)
    assert ne.total_n_data == 35820


def test_si_ridge_coefficients(si):
    _, m, train, _ = si
    say("si ridge fp64")
    ne = mlp_dev.calc_xtx_xty(m, train, precision=capi.PLM_FP64)
    coeffs = mlp_dev.solver_ridge(ne.xtx, ne.xty, np.array([1e2, 1e3]))
    np.testing.assert_allclose(coeffs[20], [0.000854, 0.000642], rtol=1e-3)
    np.testing.assert_allclose(coeffs[30], [-0.000859, -0.000607], rtol=1e-3)


def test_si_mixed_fit_predictions(si):
    """Models fitted from the mixed-precision normal equations predict the
    training snapshots within 0.1 meV/atom and 1 meV/A of the FP64 fit, for
    each alpha. (The coefficients themselves differ by up to a few percent
    along poorly determined directions, where predictions do not change.)"""
    _, m, train, _ = si
    ds = train[0]
    sts = ds.structures
    alphas = np.array([1e2, 1e3])
    fits = {}
    for prec in (capi.PLM_FP64, capi.PLM_MIXED):
        say(f"si fit prec={prec}")
        ne = mlp_dev.calc_xtx_xty(m, train, precision=prec)
        fits[prec] = mlp_dev.solver_ridge(ne.xtx, ne.xty, alphas) / ne.scales[:, None]
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    n = ds.total_n_atoms
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


def test_si_eval_accuracy(si):
    """polymlp.yaml.Si over the 180 training snapshots: energy RMSE per atom
    and force RMSE (PolymlpEvalAccuracy.compute_error_single)."""
    _, _, train, _ = si
    spec, coeffs = load_model("polymlp.yaml.Si")
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    ds = train[0]
    sts = ds.structures
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    say("si eval batch")
    E, F, _ = capi.eval_batch(m, b, capi.PLM_FP64)
    n = ds.total_n_atoms
    rmse_e = np.sqrt(np.mean(np.square((ds.energies - E) / n)))
    f_pred = np.concatenate([np.asarray(f).reshape(-1) for f in F])  # (N, 3) per structure
    rmse_f = np.sqrt(np.mean(np.square(ds.forces_flat - f_pred)))
    assert rmse_e == pytest.approx(5.84708583252534e-06)
    assert rmse_f == pytest.approx(0.0028245295894005276)


def test_si_learning_curve(si):
    """Learning curve with test = train and alphas (1e2, 1e3, 1e4)
    (tests/test_mlp_dev/test_standard_fit.py:35-49, rel 1e-2)."""
    _, m, train, _ = si
    say("si learning curve")
    log = mlp_dev.fit_learning_curve(m, train, train, np.array([1e2, 1e3, 1e4]))
    np.testing.assert_equal([entry[0] for entry in log], np.arange(1, 11) * 18)
    assert log[-3][1]["energy"] == pytest.approx(2.806378176050595e-06, rel=1e-2)
    assert log[-2][1]["energy"] == pytest.approx(2.7852516328822455e-06, rel=1e-2)
    assert log[-1][1]["energy"] == pytest.approx(2.7604438411037103e-06, rel=1e-2)
    assert log[-1][1]["stress"] is None and log[-1][1]["force"] > 0.0


def test_si_hybrid_features_x_sums(si):
    """Hybrid of the Si model with itself, type_indices [0], type_full
    (tests/test_mlp_dev/test_core_features.py:34-56, rel 1e-6)."""
    import copy

    from pypolymlp_amd import hybrid
    inp, _, train, _ = si
    spec = copy.deepcopy(inp.spec)
    spec.type_indices, spec.type_full = [0], True
    subs = [hybrid.SubModel.from_spec(spec), hybrid.SubModel.from_spec(spec)]
    say("si hybrid features_x")
    feat = hybrid.compute_features_hybrid(subs, train, capi.PLM_FP64)
    x = feat.x
    assert x.shape == (35820, 336)
    assert feat.n_data == (180, 34560, 180 * 6)
    assert tuple(feat.cumulative_n_features) == (168, 336)
    assert tuple(feat.first_indices[0]) == (0, 1260, 180)
    assert np.sum(x) == pytest.approx(10330588.900158295, rel=1e-6)
    assert np.sum(x[:, :40]) == pytest.approx(16251.02531073887, rel=1e-6)
    assert np.sum(x[:, 40:80]) == pytest.approx(401890.1325695197, rel=1e-6)
    assert np.sum(x[:, 80:120]) == pytest.approx(717054.2979810806, rel=1e-6)
    assert np.sum(x[:, 120:160]) == pytest.approx(2804300.094916472, rel=1e-6)
    assert np.sum(x[:, 160:200]) == pytest.approx(1242982.1653808185, rel=1e-6)
    assert np.sum(x[:, 200:240]) == pytest.approx(393663.7339174809, rel=1e-6)
    assert np.sum(x[:, 240:]) == pytest.approx(4754447.450082185, rel=1e-6)


def test_si_solver_cg(si):
    """Device CG on the Si normal equations: the reference's golden
    coefficients (tests/test_mlp_dev/test_gradient_solvers.py:22-28, alpha 1e2,
    gtol 1e-4, rtol 1e-2); the loop stops past iteration 100 with a true
    residual at the tolerance; the oracle's CG on the same matrix ends within
    1e-3 of the largest coefficient (the two summation orders drift over ~1000
    iterations and stop some tens of iterations apart: measured 2e-6 to 1.5e-4
    across builds whose X differs in the last bits)."""
    from oracle.numerics import solver_cg as cg_oracle
    _, m, train, _ = si
    ne = mlp_dev.calc_xtx_xty(m, train, precision=capi.PLM_FP64)
    F = len(ne.xty)
    say("si solver_cg gtol 1e-4")
    c, n_it = capi.solver_cg(ne.xtx, ne.xty, 1e2, gtol=1e-4)
    np.testing.assert_allclose([c[20], c[30]], [0.000854, -0.000859], rtol=1e-2)
    res = np.linalg.norm(ne.xty - (ne.xtx @ c + 1e2 * c)) / np.sqrt(F)
    c_ref, n_ref = cg_oracle(ne.xtx, ne.xty, 1e2, gtol=1e-4)
    say(f"  iterations: device {n_it}, oracle {n_ref}; true residual {res:.2e}")
    assert 100 < n_it < 10000 and res < 2e-4
    assert np.abs(c - c_ref).max() <= 1e-3 * np.abs(c_ref).max()
    # warm start: from the converged solution the loop stops soon after its
    # 101-iteration minimum
    c2, n2 = capi.solver_cg(ne.xtx, ne.xty, 1e2, coef0=c, gtol=1e-4)
    assert 101 <= n2 < 200
    assert np.abs(c2 - c).max() <= 1e-3 * np.abs(c).max()


def test_si_fit_cg(si):
    """fit_cg over the default alphas with test = train picks alpha 1e-3
    (tests/test_mlp_dev/test_gradient_fit.py:12-18)."""
    inp, m, train, _ = si
    say("si fit_cg")
    res = mlp_dev.fit_cg(m, train, train, inp.alphas)
    assert res.alpha == pytest.approx(0.001)
