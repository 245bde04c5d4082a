"""GPU parity of the X path against the oracle (reference C++ built from sources).

Neighbour lists: bit-exact (offset, neighbour, displacement) against
NeighborFull / NeighborHalf.

X rows (energy, stress, force) in plain FP64 against the oracle's X
(Local::gtinv_d / LocalPair::pair_d + Model row assembly): column-relative
error max|X - Xr| / max|Xr| <= 1e-10 on every column the reference resolves
(its own change under exact rotations of the structure is <= 1e-12 of the
column max, tests.helpers.oracle_x_and_noise), and <= 1e-6 of
max(column max, 1e-9 * median column max) on the remaining
cancellation-dominated columns (invariants ~1e-15 in near-symmetric cells,
which the regression zeroes), which still exposes any indexing or sign error.

Mixed precision (FP32 feature stream): see check_mixed; its accuracy claim
is made on fitted predictions (meV/atom, meV/A) in test_gpu_solve.py.

X^T X / X^T y / y^T y: normwise 1e-10 (FP64 MFMA) against numpy on the FP64
X; the mixed form (exact Gram of the FP32 X quantised to three 7-bit slices)
within 1e-5 and positive semidefinite to FP64 rounding; the int8 Gram kernel
equal to the numpy restatement of the quantised Gram to FP64 rounding.
"""

import os
import sys

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi, datasets
from tests.helpers import (load_model, oracle_x_and_noise, poscar, random_cell_structure,
                           ref_file, write_model_yaml, x_parity)

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not ref.available(), reason="oracle library not built"),
]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def colwise_rel(X, Xr):
    scale = np.max(np.abs(Xr), axis=0)
    scale[scale == 0] = 1.0
    return np.max(np.abs(X - Xr) / scale[None, :], axis=0)


def check_mixed(Xm, Xr, noise):
    """FP32 feature stream against the FP64 oracle.

    Columns the oracle resolves (rotation spread <= 1e-12 of the column max):
    column-relative error median <= 1e-5, 99th percentile <= 1e-3, max <= 1e-2.
    Cancellation-dominated columns: error within 1e3 of the FP32 analogue of
    the oracle's own FP64 spread (spread * eps32 / eps64), floored at FP32
    rounding of the column max.
    """
    colmax = np.abs(Xr).max(axis=0)
    colmax[colmax == 0] = 1.0
    err = np.abs(Xm - Xr).max(axis=0)
    spread = noise.max(axis=0)
    ok = spread <= 1e-12 * colmax
    e = err[ok] / colmax[ok]
    say(f"  mixed, resolved columns: median {np.median(e):.2e} p99 {np.quantile(e, 0.99):.2e} "
        f"max {e.max():.2e}")
    assert np.median(e) < 1e-5
    assert np.quantile(e, 0.99) < 1e-3
    assert e.max() < 1e-2
    if (~ok).any():
        eps_ratio = np.finfo(np.float32).eps / np.finfo(np.float64).eps
        bound = 1e3 * np.maximum(spread * eps_ratio, np.finfo(np.float32).eps * colmax)
        r = (err / bound)[~ok]
        say(f"  mixed, {int((~ok).sum())} cancelling columns: max err / bound {r.max():.2e}")
        assert r.max() <= 1.0


def report_worst(X, Xr, noise, k=5):
    colmax = np.abs(Xr).max(axis=0)
    colmax[colmax == 0] = 1.0
    rel = np.abs(X - Xr).max(axis=0) / colmax
    spread = noise.max(axis=0) / colmax
    for c in np.argsort(rel)[::-1][:k]:
        r = int(np.argmax(np.abs(X[:, c] - Xr[:, c])))
        say(f"    col {c}: rel err {rel[c]:.2e}, oracle spread {spread[c]:.2e}, colmax {colmax[c]:.3e}, "
            f"worst row {r}")


def _structs(sts):
    return [s.axis for s in sts], [s.positions_cartesian for s in sts], [s.types for s in sts]


def test_neighbor_full_and_half_bit_exact():
    rng = np.random.default_rng(11)
    cases = [("rocksalt", poscar("POSCAR-rocksalt"), 6.0), ("MgO", poscar("POSCAR.RS.MgO"), 8.0)]
    for k in range(12):
        ax, pos, _ = random_cell_structure(rng, n_atoms=7)
        cases.append((f"random{k}", (ax, pos), float(rng.uniform(4.0, 7.0))))
    for name, st, rc in cases:
        if isinstance(st, tuple):
            ax, pos = st
        else:
            ax, pos = st.axis, st.positions_cartesian
        for half in (False, True):
            say(f"neighbor case={name} rc={rc} half={half}")
            off, nb, tid, d = capi.neighbor(ax, pos, rc, half=half)
            ro, rn, rd = ref.neighbor_list(ax, pos, rc, half=half)
            assert np.array_equal(off, ro), name
            assert np.array_equal(nb, rn), name
            assert np.array_equal(d, rd), name


def _use_path(path, monkeypatch):
    """The linear columns are built by the signature-group kernel unless
    PLM_XROWS=generic selects the column-block kernel; PLM_XROWS=global makes
    that kernel keep its row tile in X itself. "group-queue" forces the
    signature-group kernel's persistent work-queue launch with one task per
    workgroup and "group-split" its row tile cut into two slot parts
    (PLM_XGROUP); "group" lets the autotuner choose."""
    monkeypatch.delenv("PLM_XGROUP", raising=False)
    if path in ("generic", "global"):
        monkeypatch.setenv("PLM_XROWS", path)
    else:
        monkeypatch.delenv("PLM_XROWS", raising=False)
    if path == "group-queue":
        monkeypatch.setenv("PLM_XGROUP", "1,8,1,1")
    elif path == "group-split":
        monkeypatch.setenv("PLM_XGROUP", "1,8,2,0")


def _features_x(m, b, prec, path):
    capi.profile_reset()
    capi.profile_enable(True)
    X = capi.features_x(m, b, prec)
    capi.profile_enable(False)
    n_group = capi.profile_get("xgroup")[1]
    assert (n_group > 0) == path.startswith("group"), f"path {path}: {n_group} signature-group launches"
    return X


@pytest.mark.parametrize("path", ["group", "generic"])
@pytest.mark.parametrize("model", ["polymlp.yaml.gtinv.MgO", "polymlp.yaml.pair.MgO"])
def test_x_rows_mgo(model, path, monkeypatch):
    _use_path(path, monkeypatch)
    spec, _ = load_model(model)
    m = capi.Model(spec.as_dict())
    sts = [poscar("POSCAR-00001.MgO"), poscar("POSCAR-00002.MgO")]
    axes, pos, types = _structs(sts)
    say(f"x rows model={model} path={path} FP64")
    b = capi.Batch(m, axes, pos, types)
    X = _features_x(m, b, capi.PLM_FP64, path)
    Xr, noise = oracle_x_and_noise(ref.RefModel(ref_file(model)), axes, pos, types)
    assert X.shape == Xr.shape
    strict, loose, n_loose = x_parity(X, Xr, noise)
    say(f"  fp64 column-relative err {strict:.2e} (resolved), {loose:.2e} over {n_loose} cancelling columns")
    assert strict <= 1e-10
    assert loose <= 1e-6
    say(f"x rows model={model} path={path} MIXED")
    check_mixed(_features_x(m, b, capi.PLM_MIXED, path), Xr, noise)


@pytest.mark.parametrize("path", ["group", "group-queue", "group-split", "generic", "global"])
@pytest.mark.parametrize("cfg,n_st", [("cfg1", 3), ("cfg2", 2), ("cfg3", 4)])
def test_x_rows_configs(cfg, n_st, path, tmp_path, monkeypatch):
    _use_path(path, monkeypatch)
    spec = getattr(datasets, f"config{cfg[-1]}_model")()
    if cfg == "cfg1":
        ds = datasets.config1_dataset(n_st)
    elif cfg == "cfg2":
        ds = datasets.config2_dataset(n_st, targets=False)
    else:
        ds = datasets.config3_dataset(40, targets=False)
        ds.structures = sorted(ds.structures, key=lambda s: s.n_total)[:n_st]
    m = capi.Model(spec.as_dict())
    axes, pos, types = _structs(ds.structures)
    say(f"x rows {cfg} path={path} n_st={n_st} atoms={[len(t) for t in types]}")
    b = capi.Batch(m, axes, pos, types)
    X = _features_x(m, b, capi.PLM_FP64, path)
    yaml_path = write_model_yaml(spec, str(tmp_path / "model.yaml"), np.ones(m.n_features))
    Xr, noise = oracle_x_and_noise(ref.RefModel(yaml_path), axes, pos, types)
    strict, loose, n_loose = x_parity(X, Xr, noise)
    say(f"  fp64 column-relative err {strict:.2e} (resolved), {loose:.2e} over {n_loose} cancelling columns")
    report_worst(X, Xr, noise)
    assert np.linalg.norm(X - Xr) <= 1e-10 * np.linalg.norm(Xr)
    assert strict <= 1e-10
    assert loose <= 1e-6
    check_mixed(_features_x(m, b, capi.PLM_MIXED, path), Xr, noise)


def test_normal_equations_cfg1():
    ds = datasets.config1_dataset(5)
    spec = datasets.config1_model()
    m = capi.Model(spec.as_dict())
    axes, pos, types = _structs(ds.structures)
    b = capi.Batch(m, axes, pos, types)
    say("normal equations cfg1 FP64")
    Xr = capi.features_x(m, b, capi.PLM_FP64)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.1, 1.0, Xr.shape[0])
    y = rng.normal(size=Xr.shape[0])
    acc = capi.Normal(m)
    acc.accumulate(b, w, w * y, capi.PLM_FP64)
    xtx, xty, s, s2, ysq, n = acc.result()
    Xw = Xr * w[:, None]
    ref_xtx = Xw.T @ Xw
    assert n == Xr.shape[0]
    assert np.linalg.norm(xtx - ref_xtx) / np.linalg.norm(ref_xtx) < 1e-10
    np.testing.assert_allclose(xty, Xw.T @ (w * y), rtol=1e-10, atol=1e-10 * np.abs(Xw.T @ (w * y)).max())
    ne = len(ds.structures)
    np.testing.assert_allclose(s, Xr[:ne].sum(axis=0), rtol=1e-12)
    np.testing.assert_allclose(s2, (Xr[:ne] ** 2).sum(axis=0), rtol=1e-12)
    say("normal equations cfg1 MIXED")
    accm = capi.Normal(m)
    accm.accumulate(b, w, w * y, capi.PLM_MIXED)
    xtxm, xtym, _, _, ysqm, _ = accm.result()
    rel = np.linalg.norm(xtxm - ref_xtx) / np.linalg.norm(ref_xtx)
    say(f"  mixed X^T X normwise rel err {rel:.2e}")
    assert rel < 1e-5
    ref_xty = Xw.T @ (w * y)
    assert np.linalg.norm(xtym - ref_xty) / np.linalg.norm(ref_xty) < 1e-5
    assert ysqm == pytest.approx(float((w * y) @ (w * y)), rel=1e-5)
    assert ysq == pytest.approx(float((w * y) @ (w * y)), rel=1e-12)
    # the mixed Gram of [X | y] is the exact Gram of one quantised matrix:
    # positive semidefinite to FP64 rounding (a Gram with dropped products or
    # rounded FP32 partial sums is indefinite far above this level)
    G = np.block([[xtxm, xtym[:, None]], [xtym[None, :], np.array([[ysqm]])]])
    d = 1.0 / np.sqrt(np.maximum(np.diag(G), 1e-300))
    ev = np.linalg.eigvalsh(G * d[:, None] * d[None, :])
    say(f"  mixed [X|y] Gram, unit diagonal: eigenvalue min {ev.min():.2e} max {ev.max():.2e}")
    assert ev.min() > -1e-10 * ev.max()


def test_int8_gram_is_exact_gram_of_quantised_x():
    """The int8 Gram kernel against the numpy restatement of its quantisation
    (tests.helpers.quantised_gram): through PLM_MIXED64 (FP64 X, the same
    quantisation and Gram kernels as PLM_MIXED) with unit row weights, X^T X,
    X^T y and y^T y equal the exact Gram of [X | y] quantised per (chunk,
    column), the energy and stress rows in a chunk of their own and the force
    rows in batch order in 4096-row chunks (normal.cpp's quantisation order),
    to FP64 rounding: 1e-13 of sqrt(G_ii G_jj) per entry. A dropped or
    misplaced slice product, a wrong exponent or a transposed tile is off by
    1e-8 or more of the Gram (the quantisation error itself is 5e-8). The X the
    accumulation builds can differ from plm_features_x's in the last bit (its
    row sums use atomics); a last digit that flips with it moves one row and
    column of the Gram by up to 2^-21 of its scale, so the per-entry bound holds
    on 99.9 % of the entries and the normwise difference is bounded at 1e-10."""
    from tests.helpers import quant_chunks, quantised_gram
    ds = datasets.config1_dataset(30)  # 4,860 force rows: a full and a partial chunk
    m = capi.Model(datasets.config1_model().as_dict())
    axes, pos, types = _structs(ds.structures)
    b = capi.Batch(m, axes, pos, types)
    say("int8 Gram vs quantised Gram: FP64 X of 30 cfg1 structures")
    X = capi.features_x(m, b, capi.PLM_FP64)
    rng = np.random.default_rng(11)
    y = rng.normal(size=X.shape[0])
    acc = capi.Normal(m)
    acc.accumulate(b, np.ones(X.shape[0]), y, capi.PLM_MIXED64)
    xtx, xty, _, _, ysq, _ = acc.result()
    # quantisation order: the 7 energy and stress rows of every structure, then
    # every force row in batch order
    order = b.row_order()
    ref_of_batch = np.empty_like(order)
    ref_of_batch[order] = np.arange(len(order))
    n_es = 7 * len(ds.structures)
    starts = np.concatenate([[0], np.cumsum(7 + 3 * np.asarray(b.n_atoms))])
    force_batch = np.concatenate([np.arange(s + 7, e) for s, e in zip(starts[:-1], starts[1:])])
    rows = np.concatenate([np.arange(n_es), ref_of_batch[force_batch]])
    A = np.concatenate([X, y[:, None]], axis=1)[rows]
    G = quantised_gram(A, quant_chunks(n_es, A.shape[0]))
    F = X.shape[1]
    Gk = np.empty_like(G)
    Gk[:F, :F], Gk[:F, F], Gk[F, :F], Gk[F, F] = xtx, xty, xty, ysq
    d = np.sqrt(np.maximum(np.diag(G), 1e-300))
    rel = np.abs(Gk - G) / (d[:, None] * d[None, :])
    norm = np.linalg.norm(Gk - G) / np.linalg.norm(G)
    exact = np.linalg.norm(A.T @ A - G) / np.linalg.norm(G)
    say(f"  per entry: 99.9 % {np.quantile(rel, 0.999):.1e}, max {rel.max():.1e}; "
        f"normwise {norm:.1e} (quantisation error {exact:.1e})")
    assert np.quantile(rel, 0.999) <= 1e-13
    assert norm <= 1e-10
    assert exact > 1e-9  # the comparison distinguishes quantised from exact


@pytest.mark.parametrize("prec", [capi.PLM_FP64, capi.PLM_MIXED])
def test_normal_equations_sub_batches(prec, monkeypatch):
    """One accumulate call split into several sub-batches (X budget 8 MB,
    about 7 sub-batches of cfg1): X^T X, X^T y, y^T y and the energy-row sums
    equal those of a single sub-batch: to
    rounding in FP64; in the mixed form to 1e-6, because the quantisation
    exponents are shared per 4096-row chunk of a sub-batch and the chunks move
    with the sub-batch boundaries (each Gram is exact for its own quantised X)."""
    ds = datasets.config1_dataset(40)
    m = capi.Model(datasets.config1_model().as_dict())
    axes, pos, types = _structs(ds.structures)
    b = capi.Batch(m, axes, pos, types)
    rng = np.random.default_rng(8)
    w = rng.uniform(0.1, 1.0, b.n_rows)
    y = rng.normal(size=b.n_rows)
    out = {}
    for budget in (None, "8"):
        if budget is None:
            monkeypatch.delenv("PLM_X_BUDGET_MB", raising=False)
        else:
            monkeypatch.setenv("PLM_X_BUDGET_MB", budget)
        say(f"normal equations prec={prec} x budget={budget}")
        acc = capi.Normal(m)
        acc.accumulate(b, w, w * y, prec)
        acc.accumulate(b, w, w * y, prec)
        out[budget] = acc.result()
    a, c = out[None], out["8"]
    assert a[5] == c[5] == 2 * b.n_rows
    tol = 1e-12 if prec == capi.PLM_FP64 else 1e-6
    for k in range(4):
        err = np.linalg.norm(c[k] - a[k]) / np.linalg.norm(a[k])
        say(f"  output {k}: relative difference {err:.2e}")
        assert err <= (1e-12 if k >= 2 else tol)
    assert c[4] == pytest.approx(a[4], rel=tol)


@pytest.mark.parametrize("shape", [None, "1,8,2,1", "1,8,4,0", "global"])
def test_x_rows_largest_cfg3(shape, monkeypatch):
    """The largest structure of the seeded 40-structure ternary set (429 atoms,
    1294 rows x 25,428 columns): its rows do not fit one task's LDS row tile,
    so the signature-group kernel cuts the tile into slot parts (two by
    default; four, and the persistent work queue, forced by PLM_XGROUP), or
    the column-block kernel accumulates it in X (global). FP64 against the
    oracle's seeded projections, energy row and column norms
    (tests/golden/make_cfg3_large.py); the FP32 stream against FP64
    column-relative."""
    monkeypatch.delenv("PLM_XGROUP", raising=False)
    monkeypatch.delenv("PLM_XROWS", raising=False)
    if shape == "global":
        monkeypatch.setenv("PLM_XROWS", "global")
    elif shape is not None:
        monkeypatch.setenv("PLM_XGROUP", shape)
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_cfg3_large as mk
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg3_large.npz"))
    st = mk.largest_structure()
    assert st.n_total == int(g["n_atoms"])
    m = capi.Model(datasets.config3_model().as_dict())
    say(f"x rows cfg3 largest structure, {st.n_total} atoms, shape {shape}")
    b = capi.Batch(m, [st.axis], [st.positions_cartesian], [st.types])
    capi.profile_reset()
    capi.profile_enable(True)
    X = capi.features_x(m, b, capi.PLM_FP64)
    capi.profile_enable(False)
    assert (capi.profile_get("xgroup")[1] > 0) == (shape != "global")
    assert X.shape == tuple(g["shape"])
    p, q = mk.projections(X)
    for name, a, r in (("u^T X", p, g["ut_x"]), ("X v", q, g["x_v"]), ("energy row", X[0], g["energy_row"])):
        err = np.linalg.norm(a - r) / np.linalg.norm(r)
        say(f"  fp64 {name}: relative error {err:.2e}")
        assert err <= 1e-10
    cn = np.linalg.norm(X, axis=0)
    np.testing.assert_allclose(cn, g["col_norm"], rtol=1e-10, atol=1e-10 * g["col_norm"].max())
    Xm = capi.features_x(m, b, capi.PLM_MIXED)
    colmax = np.abs(X).max(axis=0)
    nz = colmax > 0
    e = np.abs(Xm - X).max(axis=0)[nz] / colmax[nz]
    say(f"  mixed column-relative: median {np.median(e):.2e} p99 {np.quantile(e, 0.99):.2e} max {e.max():.2e}")
    assert np.median(e) < 1e-5
    assert np.quantile(e, 0.99) < 1e-3


def test_autotuner_choice_matches_rule(monkeypatch):
    """The signature-group launch shape chosen by timing (plm_model_tuning
    reports it) builds the same X as the rule-based shape (PLM_AUTOTUNE=0), up
    to the order of the FP64 atomic adds."""
    monkeypatch.delenv("PLM_XGROUP", raising=False)
    monkeypatch.delenv("PLM_XROWS", raising=False)
    ds = datasets.config2_dataset(24, seed=2, targets=False)
    axes, pos, types = _structs(ds.structures)
    say("autotuner: binary alloy, 24 structures")
    m = capi.Model(datasets.config2_model().as_dict())
    b = capi.Batch(m, axes, pos, types)
    X = capi.features_x(m, b, capi.PLM_FP64)
    rep = capi.tuning_report(m)
    say(rep)
    assert "xgroup fp64" in rep
    monkeypatch.setenv("PLM_AUTOTUNE", "0")
    m2 = capi.Model(datasets.config2_model().as_dict())
    b2 = capi.Batch(m2, axes, pos, types)
    X2 = capi.features_x(m2, b2, capi.PLM_FP64)
    assert capi.tuning_report(m2) == ""
    assert np.linalg.norm(X - X2) <= 1e-12 * np.linalg.norm(X2)
