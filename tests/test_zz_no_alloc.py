"""plm_eval_noalloc (evaluation with every device buffer in a fixed file-scope
device array, no allocation in the call) against plm_eval, through the C-ABI.

Two calls in a row, the second on a smaller structure (the array keeps the
first call's contents), in FP64 and in mixed precision, each compared with
plm_eval on the same structure: the kernels are the same and the neighbour
lists are the reference's in the same order, so energies and forces agree to
1e-12 (relative, normwise) -- tighter than the FP64 parity bound of 1e-10 --
and stresses to 1e-11 (they are per-wave partial sums added by FP64 atomics in
no fixed order, of pair terms that cancel to about 1e-3 of their size). The
FP64 energy of the reference's MgO cell also matches the reference's golden
value. Structures whose buffers exceed the array are refused with an error.
"""

import sys

import numpy as np
import pytest

from pypolymlp_amd import capi
from pypolymlp_amd.structure import supercell
from tests.helpers import load_model, poscar

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) <= tol * max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def mgo():
    spec, coeffs = load_model("polymlp.yaml.gtinv.MgO")
    model = capi.Model(spec.as_dict())
    model.set_coeffs(coeffs)
    return model


@pytest.mark.parametrize("precision", [capi.PLM_FP64, capi.PLM_MIXED])
def test_eval_noalloc_matches_eval_two_calls(mgo, precision):
    base = poscar("POSCAR.RS.MgO")
    big = supercell(base, [6, 6, 6])   # 1728 atoms: plm_eval uses its cell-list search
    rng = np.random.default_rng(5)
    pos_big = big.positions_cartesian + rng.normal(0.0, 0.05, big.positions_cartesian.shape)
    for axis, pos, types in ((big.axis, pos_big, big.types),
                             (base.axis, base.positions_cartesian, base.types)):
        n = len(types)
        say(f"noalloc: MgO gtinv, {n} atoms, precision {precision}")
        e0, f0, s0 = capi.eval_structure(mgo, axis, pos, types, precision)
        e1, f1, s1 = capi.eval_structure_noalloc(mgo, axis, pos, types, precision)
        assert abs(e1 - e0) <= 1e-12 * abs(e0)
        assert _close(f1, f0) and _close(s1, s0, 1e-11)
        if precision == capi.PLM_FP64 and n == 8:
            assert e1 == pytest.approx(-40.223320043232334, rel=1e-8)


def test_eval_noalloc_refuses_oversized_structure(mgo):
    st = supercell(poscar("POSCAR.RS.MgO"), [14, 14, 14])   # 21,952 atoms
    say(f"noalloc: oversized {len(st.types)} atoms")
    with pytest.raises(RuntimeError, match="too large"):
        capi.eval_structure_noalloc(mgo, st.axis, st.positions_cartesian, st.types)
    # the array is still usable afterwards
    base = poscar("POSCAR.RS.MgO")
    e, _, _ = capi.eval_structure_noalloc(mgo, base.axis, base.positions_cartesian, base.types)
    assert e == pytest.approx(-40.223320043232334, rel=1e-8)


def test_eval_noalloc_needs_coefficients():
    spec, _ = load_model("polymlp.yaml.gtinv.MgO")
    model = capi.Model(spec.as_dict())
    base = poscar("POSCAR.RS.MgO")
    with pytest.raises(RuntimeError, match="coefficients"):
        capi.eval_structure_noalloc(model, base.axis, base.positions_cartesian, base.types)
