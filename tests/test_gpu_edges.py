"""Edge cases of the batch path against the oracle, and input errors.

One ragged batch mixes an isolated atom (no neighbours: X and energy are
zero), one atom that sees only its own periodic images, a pair of atoms of
one type in a binary model, a 7-atom skewed random cell and the 64-atom MgO
supercell. X (both linear-column kernels, FP64 and mixed) and the batch
evaluation are compared with the reference's compiled code on the same
structures. Invalid inputs are rejected on the host with an error before any
launch.
"""

import sys

import numpy as np
import pytest

from oracle import ref
from pypolymlp_amd import capi
from tests.helpers import load_model, poscar, random_cell_structure, ref_file

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle library not built")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ragged():
    rng = np.random.default_rng(3)
    axes, pos, types = [], [], []
    axes.append(np.eye(3) * 20.0), pos.append(np.array([[1.0], [2.0], [3.0]])), types.append([0])
    axes.append(np.eye(3) * 2.9), pos.append(np.array([[0.1], [0.2], [0.3]])), types.append([1])
    axes.append(np.eye(3) * 20.0)
    pos.append(np.array([[1.0, 3.0], [2.0, 2.5], [3.0, 3.1]]))
    types.append([1, 1])
    ax, p, t = random_cell_structure(rng, n_atoms=7)
    axes.append(ax), pos.append(p), types.append(list(t))
    st = poscar("POSCAR-00001.MgO")
    axes.append(st.axis), pos.append(st.positions_cartesian), types.append(list(st.types))
    return axes, pos, types


@pytest.fixture(scope="module")
def mgo():
    spec, coeffs = load_model("polymlp.yaml.gtinv.MgO")
    m = capi.Model(spec.as_dict())
    m.set_coeffs(coeffs)
    return m, ref.RefModel(ref_file("polymlp.yaml.gtinv.MgO"), with_eval=True)


@pytest.mark.parametrize("path", ["group", "generic"])
def test_ragged_batch_x_matches_oracle(mgo, path, monkeypatch):
    if path == "generic":
        monkeypatch.setenv("PLM_XROWS", "generic")
    else:
        monkeypatch.delenv("PLM_XROWS", raising=False)
    m, r = mgo
    axes, pos, types = _ragged()
    b = capi.Batch(m, axes, pos, types)
    Xr = r.x_batch(axes, pos, types)
    say(f"ragged X path={path} atoms={[len(t) for t in types]}")
    X = capi.features_x(m, b, capi.PLM_FP64)
    assert X.shape == Xr.shape
    assert not np.any(X[0]) and not np.any(Xr[0])  # isolated atom: energy row zero
    assert np.linalg.norm(X - Xr) <= 1e-10 * np.linalg.norm(Xr)
    Xm = capi.features_x(m, b, capi.PLM_MIXED)
    assert np.linalg.norm(Xm - Xr) <= 1e-4 * np.linalg.norm(Xr)


def test_ragged_batch_eval_matches_oracle(mgo):
    m, r = mgo
    axes, pos, types = _ragged()
    b = capi.Batch(m, axes, pos, types)
    say("ragged eval")
    E, F, S = capi.eval_batch(m, b, capi.PLM_FP64)
    for k in range(len(axes)):
        e, f, s = r.eval(axes[k], pos[k], types[k])
        f = np.asarray(f).reshape(-1, 3)
        assert E[k] == pytest.approx(e, rel=1e-10, abs=1e-12)
        assert np.abs(F[k] - f).max() <= 1e-9 * max(1.0, np.abs(f).max())
        assert np.abs(S[k] - np.asarray(s)).max() <= 1e-9 * max(1.0, np.abs(s).max())
    assert E[0] == 0.0 and not np.any(F[0])


def test_invalid_inputs_are_errors(mgo):
    m, _ = mgo
    ax = np.eye(3) * 10.0
    p = np.array([[1.0, 2.0], [1.0, 2.0], [1.0, 2.0]])
    with pytest.raises(RuntimeError, match="type"):
        capi.Batch(m, [ax], [p], [[0, 2]])
    with pytest.raises(RuntimeError, match="type"):
        capi.Batch(m, [ax], [p], [[-1, 0]])
    with pytest.raises(RuntimeError):
        capi.Batch(m, [ax, ax], [p, np.zeros((3, 0))], [[0, 1], []])
    with pytest.raises((RuntimeError, ValueError)):
        capi.Batch(m, [], [], [])
    with pytest.raises(RuntimeError):
        capi.solver_cg(np.eye(3), np.ones(3), 1.0, max_iter=-1)
