"""GPU evaluation (PolymlpEval) against the oracle and the reference's golden values.

Golden: tests/test_calc/test_properties_MgO.py:9-95 (energy rel 1e-8, forces
atol 1e-6 eV/A, stress atol 1e-5 GPa). Oracle: PolymlpEval compiled from the
reference sources, energy rel 1e-10, forces and stress normwise 1e-10.
"""

import sys

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi, datasets
from pypolymlp_amd.structure import volume_gpa_factor
from tests.helpers import load_model, poscar, ref_file, write_model_yaml
from tests.test_oracle_golden import MGO_GTINV_F, MGO_GTINV_S, MGO_PAIR_F, MGO_PAIR_S

pytestmark = [pytest.mark.gpu]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# the half-pair loop's two forms (one lane per pair, pair records in LDS), the
# contraction's two workgroup sizes (terms dealt to 64 or 256 lanes) and the
# order parameters' forms (register sums of Q entries per thread over NT
# threads, QxNT, or sums in LDS); "auto" is the library's choice of all three
PAIR_FORMS = ["auto", "lane", "records", "terms64", "terms256",
              "anlm2x128", "anlm3x256", "anlm3x512", "anlm4x512", "anlmlds"]


def _pair_form(form, monkeypatch):
    for var in ("PLM_EVAL_PAIR", "PLM_PRODSUM_LANES", "PLM_ANLM_SHAPE"):
        monkeypatch.delenv(var, raising=False)
    if form in ("lane", "records"):
        monkeypatch.setenv("PLM_EVAL_PAIR", form)
    elif form.startswith("terms"):
        monkeypatch.setenv("PLM_PRODSUM_LANES", form[5:])
    elif form.startswith("anlm"):
        monkeypatch.setenv("PLM_ANLM_SHAPE", form[4:])


# heads each forced register form holds per type (Q x NT)
ANLM_CAPACITY = {"anlm2x128": 256, "anlm3x256": 768, "anlm3x512": 1536, "anlm4x512": 2048}


def _refused(form, m, n_type):
    """True where a forced order-parameter form cannot hold the model's heads:
    the evaluation must then fail with an error before any result."""
    heads = max(m.type_sizes(t)[1] for t in range(n_type))
    return heads > ANLM_CAPACITY.get(form, 1 << 30)


@pytest.mark.parametrize("form", PAIR_FORMS)
@pytest.mark.parametrize(
    "model,e_true,f_true,s_true",
    [
        ("polymlp.yaml.pair.MgO", -40.22469744315832, MGO_PAIR_F, MGO_PAIR_S),
        ("polymlp.yaml.gtinv.MgO", -40.223320043232334, MGO_GTINV_F, MGO_GTINV_S),
    ],
)
def test_eval_mgo_golden(model, e_true, f_true, s_true, form, monkeypatch):
    _pair_form(form, monkeypatch)
    spec, coeffs = load_model(model)
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    st = poscar("POSCAR.RS.MgO")
    if _refused(form, m, spec.n_type):
        say(f"eval model={model}: form {form} refused")
        with pytest.raises(RuntimeError, match="exceed"):
            capi.eval_structure(m, st.axis, st.positions_cartesian, st.types)
        return
    say(f"eval model={model} pair form={form}")
    e, f, s = capi.eval_structure(m, st.axis, st.positions_cartesian, st.types)
    assert e == pytest.approx(e_true, rel=1e-8)
    np.testing.assert_allclose(f, f_true, atol=1e-6)
    np.testing.assert_allclose(s * volume_gpa_factor(st), s_true, atol=1e-5)
    if ref.available():
        er, fr, sr = ref.RefModel(ref_file(model), with_eval=True).eval(
            st.axis, st.positions_cartesian, st.types)
        assert e == pytest.approx(er, rel=1e-10)
        assert np.linalg.norm(f - fr) <= 1e-10 * np.linalg.norm(fr)
        assert np.linalg.norm(s - sr) <= 1e-10 * np.linalg.norm(sr)


# oracle results per configuration (the same seeded coefficients in every form)
_ORACLE_EVAL = {}


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
@pytest.mark.parametrize("form", PAIR_FORMS)
@pytest.mark.parametrize("cfg", ["cfg1", "cfg2", "cfg3"])
def test_eval_configs_vs_oracle(cfg, form, tmp_path, monkeypatch):
    """cfg1 (one type), cfg2 (two types) and cfg3 (three types, 1,755 heads and
    8,268 linear features per type; its 275- and 82-atom structures) against
    the oracle's PolymlpEval in every kernel form."""
    _pair_form(form, monkeypatch)
    if cfg == "cfg1":
        spec, ds = datasets.config1_model(), datasets.config1_dataset(3)
        sts = ds.structures
    elif cfg == "cfg2":
        spec, ds = datasets.config2_model(), datasets.config2_dataset(2, targets=False)
        sts = ds.structures
    else:
        spec, ds = datasets.config3_model(), datasets.config3_dataset(3, targets=False)
        sts = ds.structures[1:]
    m = capi.Model(spec.as_dict())
    rng = np.random.default_rng(7)
    coeffs = rng.normal(size=m.n_features) * 1e-3
    m.set_coeffs(coeffs)
    if cfg not in _ORACLE_EVAL:
        path = write_model_yaml(spec, str(tmp_path / "m.yaml"), coeffs)
        R = ref.RefModel(path, with_eval=True)
        _ORACLE_EVAL[cfg] = [R.eval(st.axis, st.positions_cartesian, st.types) for st in sts]
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    if _refused(form, m, spec.n_type):
        say(f"eval batch {cfg}: form {form} refused")
        with pytest.raises(RuntimeError, match="exceed"):
            capi.eval_batch(m, b, capi.PLM_FP64)
        return
    say(f"eval batch {cfg} FP64, pair form {form}")
    E, F, S = capi.eval_batch(m, b, capi.PLM_FP64)
    say(f"eval batch {cfg} MIXED")
    Em, Fm, Sm = capi.eval_batch(m, b, capi.PLM_MIXED)
    for k, st in enumerate(sts):
        er, fr, sr = _ORACLE_EVAL[cfg][k]
        assert E[k] == pytest.approx(er, rel=1e-10)
        assert np.linalg.norm(F[k] - fr) <= 1e-10 * np.linalg.norm(fr)
        assert np.linalg.norm(S[k] - sr) <= 1e-10 * np.linalg.norm(sr)
        n = st.n_total
        say(f"  mixed: dE/atom {abs(Em[k] - er) / n * 1e3:.3e} meV, "
            f"max dF {np.abs(Fm[k] - fr).max() * 1e3:.3e} meV/A")
        assert abs(Em[k] - er) / n < 1e-4
        assert np.abs(Fm[k] - fr).max() < 1e-3


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
def test_large_cell_binned_lists_and_eval(tmp_path):
    """Cell-list search (>= 1024 atoms): bit-exact full/half lists and eval vs the oracle."""
    st = datasets.config4_structure(seed=9, reps=(8, 8, 10))  # 1280 atoms, 26.4 x 26.4 x 33 A
    ax, pos = st.axis, st.positions_cartesian
    for half in (False, True):
        say(f"binned neighbor list n={st.n_total} half={half}")
        off, nb, tid, d = capi.neighbor(ax, pos, 6.0, half=half)
        ro, rn, rd = ref.neighbor_list(ax, pos, 6.0, half=half)
        assert np.array_equal(off, ro)
        assert np.array_equal(nb, rn)
        assert np.array_equal(d, rd)
    spec = datasets.config1_model()
    m = capi.Model(spec.as_dict())
    coeffs = np.random.default_rng(3).normal(size=m.n_features) * 1e-3
    m.set_coeffs(coeffs)
    say("binned eval")
    e, f, s = capi.eval_structure(m, ax, pos, st.types)
    path = write_model_yaml(spec, str(tmp_path / "m.yaml"), coeffs)
    er, fr, sr = ref.RefModel(path, with_eval=True).eval(ax, pos, st.types)
    assert e == pytest.approx(er, rel=1e-10)
    assert np.linalg.norm(f - fr) <= 1e-10 * np.linalg.norm(fr)
    assert np.linalg.norm(s - sr) <= 1e-10 * np.linalg.norm(sr)


@pytest.mark.skipif(not ref.available(), reason="oracle library not built")
def test_eval_32000_atom_cell_vs_oracle(tmp_path):
    """The 32,000-atom bcc cell of the evaluation workload (20 x 20 x 40,
    a = 3.30 A, sigma = 0.1 A) with the 1-type order-2 max_l 4 model: the
    half neighbour list bit-exact (offsets, neighbours, differences) and FP64
    energy, forces and stress to 1e-10 (relative, normwise) against the
    reference's NeighborHalf and PolymlpEval; the mixed evaluation within
    1 meV/A and 0.1 meV/atom of FP64."""
    st = datasets.config4_structure()
    ax, pos = st.axis, st.positions_cartesian
    say(f"32000-atom half list n={st.n_total}")
    off, nb, tid, d = capi.neighbor(ax, pos, 6.0, half=True)
    ro, rn, rd = ref.neighbor_list(ax, pos, 6.0, half=True)
    assert np.array_equal(off, ro)
    assert np.array_equal(nb, rn)
    assert np.array_equal(d, rd)
    spec = datasets.config1_model()
    m = capi.Model(spec.as_dict())
    coeffs = np.random.default_rng(4).normal(size=m.n_features) * 1e-3
    m.set_coeffs(coeffs)
    say("32000-atom eval FP64 and mixed")
    e, f, s = capi.eval_structure(m, ax, pos, st.types)
    em, fm, sm = capi.eval_structure(m, ax, pos, st.types, capi.PLM_MIXED)
    path = write_model_yaml(spec, str(tmp_path / "m.yaml"), coeffs)
    er, fr, sr = ref.RefModel(path, with_eval=True).eval(ax, pos, st.types)
    assert e == pytest.approx(er, rel=1e-10)
    assert np.linalg.norm(f - fr) <= 1e-10 * np.linalg.norm(fr)
    assert np.linalg.norm(s - sr) <= 1e-10 * np.linalg.norm(sr)
    assert abs(em - e) / st.n_total < 1e-4
    assert np.abs(fm - f).max() < 1e-3


# --- legacy-format models, hybrid "flexible" models, skewed cells ------------
from tests.test_oracle_golden import MGO_HNF, MGO_HNF_ENERGY, SRTIO3_LEGACY  # noqa: E402


@pytest.mark.parametrize("name,e_true,s_true", SRTIO3_LEGACY)
def test_eval_legacy_srtio3_golden(name, e_true, s_true):
    """polymlp.lammps models read by the legacy reader
    (tests/test_calc/test_properties_legacy_SrTiO3.py:19-69)."""
    from pypolymlp_amd.params import load_mlp
    spec, coeffs = load_mlp(ref_file(name))
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    st = poscar("POSCAR.perovskite.SrTiO3")
    say(f"eval legacy model={name}")
    e, f, s = capi.eval_structure(m, st.axis, st.positions_cartesian, st.types)
    assert e == pytest.approx(e_true, abs=1e-12)
    assert f[0][0] == pytest.approx(0.0, abs=1e-12)
    np.testing.assert_allclose(s[:3] * volume_gpa_factor(st), [s_true] * 3, atol=1e-5)
    np.testing.assert_allclose(s[3:] * volume_gpa_factor(st), 0.0, atol=1e-5)


def test_eval_flexible_hybrid_srtio3_golden():
    """Hybrid of a full model and an O-only model
    (test_properties_legacy_SrTiO3.py:72-81, rel 1e-8)."""
    from pypolymlp_amd import hybrid
    from pypolymlp_amd.params import load_mlp
    subs = [hybrid.SubModel.from_spec(*load_mlp(ref_file(n)))
            for n in ("polymlp.yaml.flexible.1.SrTiO3", "polymlp.yaml.flexible.2.SrTiO3")]
    assert not subs[1].spec.type_full
    st = poscar("POSCAR.perovskite.SrTiO3")
    say("eval flexible hybrid")
    E, F, _ = hybrid.eval_hybrid(subs, [st], capi.PLM_FP64)
    assert E[0] == pytest.approx(-31.642657299368572, rel=1e-8)
    assert F[0][0][0] == pytest.approx(0.0, abs=1e-12)


def test_eval_skewed_cells_mgo_golden():
    """25 shears of the ideal rocksalt cell in one batch: every energy equals
    the reference's (tests/test_calc/test_check_neighbors.py:38-73, atol 1e-12)."""
    from pypolymlp_amd.structure import supercell
    spec, coeffs = load_model("polymlp.yaml.pair.MgO")
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    st = poscar("POSCAR.RS.idealMgO")
    sts = [supercell(st, h) for h in MGO_HNF]
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    say(f"eval skewed cells n={len(sts)}")
    E, _, _ = capi.eval_batch(m, b, capi.PLM_FP64)
    np.testing.assert_allclose(E, MGO_HNF_ENERGY, atol=1e-12)


@pytest.mark.parametrize("prec", [capi.PLM_FP64, capi.PLM_MIXED])
def test_eval_graph_replays(prec):
    """The evaluation captured as a HIP graph and replayed (1, 2 and 3 times per
    run, with eager evaluations and synchronisations in between) gives the
    eager evaluation's energies, forces and stresses on every run, and the
    golden MgO energy; a coefficient change after capture is refused."""
    spec, coeffs = load_model("polymlp.yaml.gtinv.MgO")
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    sts = [poscar("POSCAR.RS.MgO"), poscar("POSCAR.RS.idealMgO")]
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    say(f"eval graph prec={prec}")
    e0, f0, s0 = capi.eval_batch(m, b, prec)
    g = capi.EvalGraph(m, b, prec)
    for n in (1, 2, 3):
        e, f, s = g.run(n)
        say(f"  run({n}) energies {e}")
        np.testing.assert_allclose(e, e0, rtol=1e-12)
        for fk, f0k in zip(f, f0):
            np.testing.assert_allclose(fk, f0k, rtol=0, atol=1e-12 * max(1.0, np.abs(f0k).max()))
        # stresses are per-wave partial sums added by FP64 atomics in no fixed
        # order, and the pair terms cancel to about 1e-3 of their size: runs
        # differ by up to about 1e-12 of the largest component
        np.testing.assert_allclose(s, s0, rtol=0, atol=1e-11 * np.abs(s0).max())
        e1, _, _ = capi.eval_batch(m, b, prec)
        np.testing.assert_allclose(e1, e0, rtol=1e-12)
    assert e[0] == pytest.approx(-40.223320043232334, rel=1e-8 if prec == capi.PLM_FP64 else 1e-6)
    m.set_coeffs(coeffs)
    with pytest.raises(RuntimeError, match="coefficients changed"):
        g.run(1)
