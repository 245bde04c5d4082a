"""Hybrid models on the GPU: a full MgO model plus a pair model acting on the
O atoms only (type_full False, type_indices [1]).

Checked against the oracle (the reference's compiled Features / PolymlpEval):
the hybrid X equals the full model's X in its columns; in the O-only model's
columns the energy and stress rows equal the oracle's rows of the O-only
substructure, the force rows of Mg atoms are zero and those of O atoms equal
the oracle's force rows in order (py_hybrid_model.cpp:60-103). The hybrid
energy, forces and stress equal the sum of the two oracle evaluations with
the O-only forces placed on the O atoms (properties_single.py:43-79).
"""

import sys

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi, hybrid
from pypolymlp_amd.datasets import Dataset
from pypolymlp_amd.params import ModelSpec
from pypolymlp_amd.structure import PolymlpStructure
from tests.helpers import load_model, poscar, random_cell_structure, write_model_yaml

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle library not built")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cell(ax, pos, types):
    types = [int(t) for t in types]
    return PolymlpStructure(axis=ax, positions=np.linalg.solve(ax, pos),
                            n_atoms=[types.count(0), types.count(1)],
                            elements=[["Mg", "O"][t] for t in types], types=types)


def _structures():
    rng = np.random.default_rng(11)
    sts = [poscar("POSCAR-00001.MgO"), poscar("POSCAR-00002.MgO"), poscar("POSCAR.RS.MgO")]
    sts.append(_cell(*random_cell_structure(rng, n_atoms=9, n_type=2)))
    # a structure without O atoms: the O-only model contributes nothing
    ax, pos, _ = random_cell_structure(rng, n_atoms=5, n_type=2)
    sts.append(_cell(ax, pos, [0] * 5))
    return sts


def _o_only(st):
    sel = np.array([i for i, t in enumerate(st.types) if t == 1], dtype=int)
    return sel, st.positions_cartesian[:, sel], np.zeros(len(sel), dtype=int)


@pytest.fixture(scope="module")
def models(tmp_path_factory):
    spec1, c1 = load_model("polymlp.yaml.gtinv.MgO")
    spec2 = ModelSpec(n_type=1, cutoff=6.0, feature_type="pair", model_type=2, max_p=2,
                      gaussian_params2=(0.0, 5.0, 6), elements=["O"], type_full=False,
                      type_indices=[1])
    m2 = capi.Model(spec2.as_dict())
    c2 = np.random.default_rng(5).normal(size=m2.n_features) * 1e-2
    d = tmp_path_factory.mktemp("hyb")
    y1 = write_model_yaml(spec1, str(d / "m1.yaml"), c1)
    y2 = write_model_yaml(spec2, str(d / "m2.yaml"), c2)
    subs = [hybrid.SubModel.from_spec(spec1, c1), hybrid.SubModel(spec=spec2, model=m2)]
    subs[1].model.set_coeffs(c2)
    return subs, y1, y2


def test_hybrid_x_matches_oracle_composition(models):
    subs, y1, y2 = models
    sts = _structures()
    n = len(sts)
    ds = Dataset(structures=sts, energies=np.zeros(n),
                 forces=[np.zeros((3, s.n_total)) for s in sts], stresses=np.zeros((n, 6)))
    say(f"hybrid features_x n_st={n}")
    feat = hybrid.compute_features_hybrid(subs, [ds], capi.PLM_FP64)
    F1 = subs[0].model.n_features
    assert feat.cumulative_n_features == [F1, F1 + subs[1].model.n_features]
    X = feat.x
    r1, r2 = ref.RefModel(y1), ref.RefModel(y2)
    X1 = r1.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                    [s.types for s in sts])
    assert np.linalg.norm(X[:, :F1] - X1) <= 1e-10 * np.linalg.norm(X1)
    _, fb, sb = feat.first_indices[0]
    f0 = fb
    for k, st in enumerate(sts):
        sel, pos, types = _o_only(st)
        xe = X[k, F1:]
        xs = X[sb + 6 * k:sb + 6 * k + 6, F1:]
        xf = X[f0:f0 + 3 * st.n_total, F1:].reshape(st.n_total, 3, -1)
        f0 += 3 * st.n_total
        if len(sel) == 0:
            assert not np.any(xe) and not np.any(xs) and not np.any(xf)
            continue
        re, rs, rf = r2.structure_rows(st.axis, pos, types)
        scale = np.abs(rf).max() + np.abs(re).max()
        assert np.abs(xe - re).max() <= 1e-10 * scale
        assert np.abs(xs - rs).max() <= 1e-10 * scale
        mg = np.setdiff1d(np.arange(st.n_total), sel)
        assert not np.any(xf[mg])
        assert np.abs(xf[sel].reshape(-1, xf.shape[2]) - rf).max() <= 1e-10 * scale


def test_hybrid_eval_matches_oracle_sum(models):
    subs, y1, y2 = models
    sts = _structures()
    say(f"hybrid eval n_st={len(sts)}")
    E, F, S = hybrid.eval_hybrid(subs, sts, capi.PLM_FP64)
    r1, r2 = ref.RefModel(y1, with_eval=True), ref.RefModel(y2, with_eval=True)
    for k, st in enumerate(sts):
        e, f, s = r1.eval(st.axis, st.positions_cartesian, st.types)
        f = np.array(f, dtype=float).reshape(st.n_total, 3)
        s = np.array(s, dtype=float)
        sel, pos, types = _o_only(st)
        if len(sel):
            e2, f2, s2 = r2.eval(st.axis, pos, types)
            e += e2
            f[sel] += np.array(f2).reshape(len(sel), 3)
            s += np.array(s2)
        assert E[k] == pytest.approx(e, rel=1e-10)
        assert np.abs(F[k] - f).max() <= 1e-9 * max(1.0, np.abs(f).max())
        assert np.abs(S[k] - s).max() <= 1e-9 * max(1.0, np.abs(s).max())


def test_pybind_mirror_matches_host_paths(models):
    """The pybind11 mirror of libmlpcpp (the drop-in the reference's
    features.py binds): PotentialModel and PotentialHybridModel give the X of
    the host paths, with the reference's getters; PotentialPropertiesFast gives
    the same energies as the batch evaluation."""
    from pypolymlp_amd import libmlpcpp, mlp_dev
    subs, _, _ = models
    sts = _structures()
    n = len(sts)
    ds = Dataset(structures=sts, energies=np.zeros(n),
                 forces=[np.zeros((3, s.n_total)) for s in sts], stresses=np.zeros((n, 6)))
    axes = [s.axis.tolist() for s in sts]
    pos = [s.positions_cartesian.tolist() for s in sts]
    types = [list(map(int, s.types)) for s in sts]
    nat = [s.n_total for s in sts]
    say("pybind PotentialModel")
    pm = libmlpcpp.PotentialModel(subs[0].spec.as_dict(), axes, pos, types, [n], [True], nat)
    ref_x = mlp_dev.compute_features(subs[0].model, [ds], capi.PLM_FP64)
    # the X kernels sum with atomics, so two runs agree to rounding, not bitwise
    px = np.asarray(pm.get_x())
    assert px.shape == ref_x.x.shape
    assert np.linalg.norm(px - ref_x.x) <= 1e-12 * np.linalg.norm(ref_x.x)
    assert list(pm.get_n_data()) == list(ref_x.n_data)
    assert (pm.get_fbegin()[0], pm.get_sbegin()[0]) == ref_x.first_indices[0][1:]
    say("pybind PotentialHybridModel")
    hm = libmlpcpp.PotentialHybridModel([s.spec.as_dict() for s in subs], axes, pos, types,
                                        [n], [True], nat)
    hx = hybrid.compute_features_hybrid(subs, [ds], capi.PLM_FP64)
    hxx = np.asarray(hm.get_x())
    assert hxx.shape == hx.x.shape
    assert np.linalg.norm(hxx - hx.x) <= 1e-12 * np.linalg.norm(hx.x)
    assert list(hm.get_cumulative_n_features()) == hx.cumulative_n_features
    assert list(hm.get_n_data()) == list(hx.n_data)
    say("pybind PotentialPropertiesFast")
    _, c1 = load_model("polymlp.yaml.gtinv.MgO")
    pp = libmlpcpp.PotentialPropertiesFast(subs[0].spec.as_dict(), c1.tolist())
    pp.eval_multiple(axes, pos, types)
    E, _, _ = capi.eval_batch(subs[0].model,
                              capi.Batch(subs[0].model, [s.axis for s in sts],
                                         [s.positions_cartesian for s in sts], types),
                              capi.PLM_FP64)
    np.testing.assert_allclose(np.asarray(pp.get_e_array()), E, rtol=1e-12)
