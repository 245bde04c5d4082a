"""Why the mixed Gram computes every slice product (CPU).

The mixed form quantises each column of [X w | y w] per row chunk (at most
4096 rows, the energy and stress rows in a chunk of their own) to a shared
exponent E and three signed 7-bit digits, x~ = 2^(E-7) (d0 + d1 2^-7 + d2 2^-14),
the rule of kernels_gram.hip::quantize_kernel, and forms the Gram from the
int8 digit products on MFMA. This test restates that rule in numpy on the
oracle X of 40 structures of the bcc workload (6,760 rows, 2,277 columns) and
compares four Grams after the reference's column scaling
(src/pypolymlp/mlp_dev/core/utils_scales.py:6-40; the ridge adds alpha in
1e-5 .. 1e-2 to this matrix):

- the exact Gram of x~ (all nine digit products, exact integer sums): positive
  semidefinite to FP64 rounding, 5e-8 from the FP64 Gram;
- the Gram without the digit products of weight 2^-21 and 2^-28 (six products):
  eigenvalues near -1.7, so the ridge's Cholesky fails at every alpha;
- the same without only the 2^-28 product (eight products): near -6e-3;
- split bf16 (hi hi + hi lo + lo hi) with FP32 sums per chunk: near -1.3.

The largest eigenvalue is 6.3e6: perturbations of 1e-7 relative are already
far larger than alpha.
"""
import numpy as np
import pytest

from oracle import ref
from tests.helpers import quant_chunks, quant_digits, quantised_gram

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")

N_ST = 40


@pytest.fixture(scope="module")
def scaled_problem(tmp_path_factory):
    from pypolymlp_amd import datasets, mlp_dev
    from tests.helpers import write_model_yaml
    ds = datasets.config1_dataset(N_ST)
    path = str(tmp_path_factory.mktemp("gram") / "cfg1.yaml")
    rm = ref.RefModel(write_model_yaml(datasets.config1_model(), path))
    sts = ds.structures
    X = rm.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    min_e = mlp_dev.get_min_energy([ds])
    w, yw = mlp_dev.row_weights(ds, min_e)
    scales, _ = mlp_dev.round_scales(np.std(X[:N_ST], axis=0))
    A = np.concatenate([X * w[:, None], yw[:, None]], axis=1)
    # rows in the reference layout: N_ST energy rows, 6 N_ST stress rows, then
    # the force rows in the kernel's quantisation order
    bounds = quant_chunks(7 * N_ST, A.shape[0])
    return A, np.concatenate([scales, [1.0]]), bounds


def _gram_bf16x3(A, bounds):
    def bf16(x32):
        u = x32.view(np.uint32)
        u = ((u + np.uint32(0x7FFF) + ((u >> 16) & 1)) & np.uint32(0xFFFF0000)).astype(np.uint32)
        return u.view(np.float32)
    G = np.zeros((A.shape[1], A.shape[1]))
    for a, b in bounds:
        blk = A[a:b].astype(np.float32)
        hi = bf16(blk)
        lo = bf16(blk - hi)
        G += (hi.T @ hi + hi.T @ lo + lo.T @ hi).astype(np.float32).astype(np.float64)
    return G


def _scaled_min_max(G, D):
    Gs = (G / D[:, None] / D[None, :])[:-1, :-1]
    ev = np.linalg.eigvalsh(Gs)
    return ev.min(), ev.max(), Gs


def test_digits_reconstruct_21_bits():
    rng = np.random.default_rng(0)
    x = rng.normal(size=2000) * 10.0 ** rng.uniform(-6, 2, size=2000)
    e = np.frexp(np.abs(x).max())[1]
    t0, t1, t2 = quant_digits(x, e)
    assert max(np.abs(t0).max(), np.abs(t1).max(), np.abs(t2).max()) <= 127
    xq = np.ldexp(t0 + t1 / 128 + t2 / 16384, e - 7)
    # half a unit of the last digit, a whole unit where it rounds to 128 and is clipped
    err = np.abs(xq - x)
    assert err.max() <= 2.0 ** (e - 21)
    assert np.quantile(err, 0.99) <= 2.0 ** (e - 22)


def test_exact_quantised_gram_is_psd_and_truncated_grams_are_not(scaled_problem):
    A, D, bounds = scaled_problem
    ge = A.T @ A
    lam_e_min, lam_max, _ = _scaled_min_max(ge, D)
    assert lam_max > 1e6 and abs(lam_e_min) < 1e-14 * lam_max
    gq = quantised_gram(A, bounds, 4)
    assert np.linalg.norm(gq - ge) / np.linalg.norm(ge) < 1e-7
    lam_q, _, gs_q = _scaled_min_max(gq, D)
    assert lam_q > -1e-15 * lam_max
    np.linalg.cholesky(gs_q + 1e-5 * np.eye(len(gs_q)))  # every alpha of the grid
    for g in (quantised_gram(A, bounds, 2), _gram_bf16x3(A, bounds)):
        lam, _, gs = _scaled_min_max(g, D)
        assert lam < -0.1
        with pytest.raises(np.linalg.LinAlgError):
            np.linalg.cholesky(gs + 1e-2 * np.eye(len(gs)))
    lam8, _, gs8 = _scaled_min_max(quantised_gram(A, bounds, 3), D)
    assert lam8 < -1e-3
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.cholesky(gs8 + 1e-4 * np.eye(len(gs8)))
