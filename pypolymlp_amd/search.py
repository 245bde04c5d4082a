"""Surrogate-guided hyperparameter search over polymlp models.

Mirrors the reference's grid-search harness and replaces "fit every grid
point" by "fit a seeded sample, train the surrogate network on those fits,
fit only the candidates the network predicts near the Pareto front":

  grid of model candidates   src/pypolymlp/utils/grid_search/grid_utils.py:42-140,
                             grid_enum.py:275-299
  cost (ms per atom per step) src/pypolymlp/postproc/count_time.py:15-181
                             (here: GPU evaluation of copies of a 256-atom cell)
  optimal models             src/pypolymlp/utils/grid_optimal.py:10-68
                             (convex hull of (cost, energy RMSE); vertices
                             no other vertex beats in cost, energy and
                             force RMSE; energy RMSE < 30 meV/atom)

The surrogate (FP32 MLP 5-64-64-2, SiLU; pypolymlp_amd.capi.surrogate_*)
maps (cutoff, Gaussians, max_l, polynomial order, log10 F) to (log10 RMSE,
log10 cost). It is trained on the GPU by the library's own forward/backward
kernels. The exhaustive grid stays available as the fallback and as the judge
of the network's picks.
"""

from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import capi, mlp_dev
from .datasets import Dataset, _fcc_supercell, _make
from .params import ModelSpec

CUTOFFS = (5.0, 6.0, 7.0, 8.0)
GAUSSIANS = (6, 10, 14, 18)
MAX_LS = (4, 6, 8, 10)
MAX_PS = (2, 3, 4)
RMSE_THRESHOLD = 30.0  # meV/atom (grid_optimal.py:53)


@dataclass(frozen=True)
class Candidate:
    cutoff: float
    n_gauss: int
    max_l: int
    max_p: int

    def spec(self) -> ModelSpec:
        """Binary gtinv order-3 model, max_l -> gtinv max_l (L, min(L, 4)),
        Gaussian centres 0 .. cutoff - 1, model_type 3."""
        return ModelSpec(n_type=2, cutoff=self.cutoff, model_type=3, max_p=self.max_p,
                         gtinv_order=3, gtinv_maxl=(self.max_l, min(self.max_l, 4)),
                         gaussian_params1=(1.0, 1.0, 1),
                         gaussian_params2=(0.0, self.cutoff - 1.0, self.n_gauss),
                         elements=["A", "B"])

    @property
    def key(self) -> str:
        return f"rc{self.cutoff:g}_g{self.n_gauss}_l{self.max_l}_p{self.max_p}"


def grid(n: int = 160, seed: int = 5) -> List[Candidate]:
    """The 4 x 4 x 4 x 3 = 192-point grid, subsampled to n by seed."""
    full = [Candidate(*c) for c in itertools.product(CUTOFFS, GAUSSIANS, MAX_LS, MAX_PS)]
    if n >= len(full):
        return full
    idx = np.sort(np.random.default_rng(seed).choice(len(full), n, replace=False))
    return [full[i] for i in idx]


@dataclass
class Result:
    cand: Candidate
    ok: bool
    n_features: int = 0
    rmse_e: float = np.inf     # meV/atom on the test set
    rmse_f: float = np.inf     # eV/A on the test set
    cost: float = np.inf       # ms per atom per evaluation step
    seconds: float = 0.0       # wall time of fit + error + cost
    error: str = ""


def feature_count(c: Candidate) -> Optional[int]:
    """Number of X columns, or None when the reference rejects the model
    (polynomial order 4 exits in the reference's model setup)."""
    try:
        return capi.Model(c.spec().as_dict()).n_features
    except RuntimeError:
        return None


def cost_cell(n_copies: int = 50, seed: int = 11):
    """Copies of a 256-atom binary fcc cell (4 x 4 x 4, a = 3.9 A, sigma 0.05 A)."""
    rng = np.random.default_rng(seed)
    axis, frac = _fcc_supercell(3.9, (4, 4, 4))
    types = (np.arange(frac.shape[1]) % 2).astype(int)
    st = _make(axis, frac, types, ["A", "B"], rng, 0.05)
    return [st] * n_copies


def evaluate(c: Candidate, train: Dataset, test: Dataset, alphas=(1e-4, 1e-3, 1e-2),
             precision: int = capi.PLM_MIXED, cost_structures=None) -> Result:
    """Fit one candidate on the GPU, then its test errors and evaluation cost."""
    t0 = time.perf_counter()
    try:
        model = capi.Model(c.spec().as_dict())
    except RuntimeError as e:
        return Result(c, False, error=str(e), seconds=time.perf_counter() - t0)
    fr = mlp_dev.fit(model, [train], [test], np.asarray(alphas), precision=precision)
    model.set_coeffs(fr.coeffs_unscaled)
    sts = test.structures
    b = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    e, f, _ = capi.eval_batch(model, b, capi.PLM_FP64)
    n_at = np.array([s.n_total for s in sts])
    rmse_e = float(np.sqrt(np.mean(((e - test.energies) / n_at) ** 2)) * 1e3)
    df = np.concatenate([(fi - fr_.T).ravel() for fi, fr_ in zip(f, test.forces)])
    rmse_f = float(np.sqrt(np.mean(df ** 2)))
    cs = cost_structures if cost_structures is not None else cost_cell()
    cb = capi.Batch(model, [s.axis for s in cs], [s.positions_cartesian for s in cs],
                    [s.types for s in cs])
    capi.eval_batch(model, cb, capi.PLM_FP64)
    reps = 3
    tc = time.perf_counter()
    for _ in range(reps):
        capi.eval_batch(model, cb, capi.PLM_FP64)
    cost = (time.perf_counter() - tc) / reps / sum(s.n_total for s in cs) * 1e3
    return Result(c, True, model.n_features, rmse_e, rmse_f, cost, time.perf_counter() - t0)


def find_optimal(cost, rmse_e, rmse_f, use_force: bool = False, use_logscale_time: bool = False,
                 threshold: float = RMSE_THRESHOLD) -> np.ndarray:
    """Indices of the optimal models by the reference's rule
    (_find_convex_mlps / _find_lower_convex, utils/grid_optimal.py:32-68), in
    increasing cost:
      1. vertices of the convex hull of (cost, energy RMSE), or (cost, force
         RMSE) with use_force, cost on a log10 scale with use_logscale_time;
      2. of those, the ones no other vertex beats strictly in cost, energy
         RMSE and force RMSE at once;
      3. energy RMSE below the threshold (30 meV/atom).
    Fewer than three points, or collinear ones, have no hull: every point is
    then a vertex."""
    from scipy.spatial import ConvexHull

    cost = np.asarray(cost, dtype=float)
    re = np.asarray(rmse_e, dtype=float)
    rf = np.asarray(rmse_f, dtype=float)
    order = np.argsort(cost, kind="stable")  # the reference sorts its table by cost
    c, e, f = cost[order], re[order], rf[order]
    tgt = np.column_stack([np.log10(c) if use_logscale_time else c, f if use_force else e])
    try:
        verts = np.unique(ConvexHull(tgt).simplices) if len(c) >= 3 else np.arange(len(c))
    except Exception:  # degenerate (collinear) point sets
        verts = np.arange(len(c))
    keep = [i for i in verts
            if not any(j != i and c[j] < c[i] and e[j] < e[i] and f[j] < f[i] for j in verts)]
    keep = np.array([i for i in keep if e[i] < threshold], dtype=int)
    return order[keep]


def pareto(points: np.ndarray, threshold: float = RMSE_THRESHOLD) -> np.ndarray:
    """find_optimal on rows (cost, energy RMSE, force RMSE)."""
    pts = np.asarray(points, dtype=float)
    if not len(pts):
        return np.zeros(0, dtype=int)
    return find_optimal(pts[:, 0], pts[:, 1], pts[:, 2], threshold=threshold)


def hypervolume(points: np.ndarray, ref: Sequence[float]) -> float:
    """Area dominated by the points in (log10 cost, log10 rmse) up to ref."""
    p = np.log10(np.asarray(points, dtype=float))
    p = p[(p[:, 0] < ref[0]) & (p[:, 1] < ref[1])]
    if not len(p):
        return 0.0
    p = p[np.argsort(p[:, 0])]
    area, best = 0.0, ref[1]
    for k in range(len(p)):
        best = min(best, p[k, 1])
        x1 = p[k + 1, 0] if k + 1 < len(p) else ref[0]
        area += (x1 - p[k, 0]) * (ref[1] - best)
    return area


def descriptor(c: Candidate, n_features: int) -> np.ndarray:
    return np.array([c.cutoff, c.n_gauss, c.max_l, c.max_p, np.log10(max(n_features, 1))])


@dataclass
class Surrogate:
    """Standardised inputs/targets and the FP32 network parameters."""
    params: np.ndarray
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray
    loss: float = 0.0
    history: List[float] = field(default_factory=list)

    def predict(self, X: np.ndarray) -> np.ndarray:
        Z = (np.asarray(X) - self.x_mean) / self.x_std
        return capi.surrogate_predict(self.params, Z) * self.y_std + self.y_mean


def train_surrogate(X: np.ndarray, Y: np.ndarray, steps: int = 2000, lr: float = 3e-3,
                    weight_decay: float = 1e-2, seed: int = 0, chunk: int = 250) -> Surrogate:
    """Full-batch AdamW on the GPU (hand-written HIP forward/backward); the
    history starts with the loss of the initial parameters."""
    X = np.asarray(X, dtype=float)
    Y = np.asarray(Y, dtype=float)
    xm, xs = X.mean(0), X.std(0) + 1e-12
    ym, ys = Y.mean(0), Y.std(0) + 1e-12
    rng = np.random.default_rng(seed)
    n_par = capi.surrogate_n_params()
    sizes = [(64, 5), (64,), (64, 64), (64,), (2, 64), (2,)]
    parts = []
    for shp in sizes:
        if len(shp) == 2:
            parts.append(rng.normal(0.0, 1.0 / np.sqrt(shp[1]), size=shp).ravel())
        else:
            parts.append(np.zeros(shp))
    p = np.concatenate(parts).astype(np.float32)
    assert len(p) == n_par
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    Z = (X - xm) / xs
    T = (Y - ym) / ys
    hist = [capi.surrogate_grad(p, Z, T)[1]]
    t = 0
    while t < steps:
        k = min(chunk, steps - t)
        hist.append(capi.surrogate_train(p, m, v, t, Z, T, k, lr, weight_decay))
        t += k
    return Surrogate(p, xm, xs, ym, ys, hist[-1], hist)


def _targets(results: Sequence[Result]) -> np.ndarray:
    return np.array([[np.log10(r.rmse_e), np.log10(r.cost)] for r in results])


def surrogate_pick(valid: Sequence[Candidate], n_feat: Dict[Candidate, int],
                   done: Sequence[Result], n_pick: int, surr: Surrogate) -> List[Candidate]:
    """Remaining candidates closest (in predicted log RMSE) to the lower
    envelope of the fitted points and the predictions."""
    fitted = {r.cand for r in done}
    rest = [c for c in valid if c not in fitted]
    if not rest:
        return []
    pred = surr.predict(np.array([descriptor(c, n_feat[c]) for c in rest]))
    pts = np.vstack([np.column_stack([np.log10([r.cost for r in done if r.ok]),
                                      np.log10([r.rmse_e for r in done if r.ok])]),
                     pred[:, ::-1]])
    order = np.argsort(pts[:, 0])
    # lower envelope (running minimum of log RMSE over increasing cost)
    env = np.minimum.accumulate(pts[order, 1])
    x_sorted = pts[order, 0]
    gap = []
    for k, c in enumerate(rest):
        lc, lr = pred[k, 1], pred[k, 0]
        j = np.searchsorted(x_sorted, lc, side="right") - 1
        floor = env[j] if j >= 0 else np.inf
        gap.append(lr - floor)
    return [rest[i] for i in np.argsort(gap)[:n_pick]]


@dataclass
class SearchReport:
    exhaustive: List[Result]
    guided: List[Result]
    random_pick: List[Result]
    surrogate: Surrogate
    seconds_exhaustive: float
    seconds_guided: float
    front_exhaustive: List[str]
    front_guided: List[str]
    front_random: List[str]
    hv_exhaustive: float
    hv_guided: float
    hv_random: float


def run_search(cands: Sequence[Candidate], evaluate_fn: Callable[[Candidate], Result],
               n_init: int = 32, n_pick: int = 24, seed: int = 7, surrogate_steps: int = 2000,
               cached: Optional[Dict[Candidate, Result]] = None) -> SearchReport:
    """Exhaustive grid (the judge), surrogate-guided search and a random
    search with the same number of fits. Fits are deterministic, so each
    candidate is fitted once and the three searches share the results; the
    guided search's wall time is the sum of its own fits plus the network."""
    n_feat = {c: feature_count(c) for c in cands}
    valid = [c for c in cands if n_feat[c] is not None]
    res = dict(cached or {})
    t0 = time.perf_counter()
    for c in cands:
        if c not in res:
            res[c] = evaluate_fn(c) if n_feat[c] is not None else Result(c, False, error="rejected")
    seconds_ex = sum(r.seconds for r in res.values())
    rng = np.random.default_rng(seed)
    init = [valid[i] for i in rng.choice(len(valid), min(n_init, len(valid)), replace=False)]
    done = [res[c] for c in init]
    ts = time.perf_counter()
    ok = [r for r in done if r.ok]
    surr = train_surrogate(np.array([descriptor(r.cand, r.n_features) for r in ok]), _targets(ok),
                           steps=surrogate_steps, seed=seed)
    picks = surrogate_pick(valid, n_feat, done, n_pick, surr)
    t_surr = time.perf_counter() - ts
    guided = done + [res[c] for c in picks]
    seconds_guided = sum(r.seconds for r in guided) + t_surr
    rest = [c for c in valid if c not in set(init)]
    rnd = [rest[i] for i in rng.choice(len(rest), min(n_pick, len(rest)), replace=False)]
    random_pick = done + [res[c] for c in rnd]
    exhaustive = [res[c] for c in cands]

    def front(rs):
        good = [r for r in rs if r.ok]
        if not good:
            return [], []
        pts = np.array([[r.cost, r.rmse_e, r.rmse_f] for r in good])
        idx = pareto(pts)
        return [good[i].cand.key for i in idx], pts[idx][:, :2]

    fe, pe = front(exhaustive)
    fg, pg = front(guided)
    fr, pr = front(random_pick)
    allpts = np.array([[r.cost, r.rmse_e] for r in exhaustive if r.ok])
    ref = (np.log10(allpts[:, 0].max()) + 0.05, np.log10(RMSE_THRESHOLD))
    del t0
    return SearchReport(exhaustive, guided, random_pick, surr, seconds_ex, seconds_guided,
                        fe, fg, fr, hypervolume(pe, ref) if len(pe) else 0.0,
                        hypervolume(pg, ref) if len(pg) else 0.0,
                        hypervolume(pr, ref) if len(pr) else 0.0)


def search_data(n_structures: int = 250, n_test: int = 50):
    """(train, test) of the search: the first n_structures - n_test seeded
    binary-alloy structures (128-atom fcc 4x4x2 cells) for training, the last
    n_test for the errors."""
    from .datasets import config2_dataset

    ds = config2_dataset(n_structures)
    n_tr = n_structures - n_test

    def sub(i0, i1, name):
        return Dataset(structures=ds.structures[i0:i1], energies=ds.energies[i0:i1],
                       forces=ds.forces[i0:i1], stresses=ds.stresses[i0:i1], name=name)
    return sub(0, n_tr, "train"), sub(n_tr, n_structures, "test")


def main(argv=None):
    """python -m pypolymlp_amd.search --out report.json: the cfg5 search on a
    seeded subset of the binary-alloy dataset (train/test split)."""
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("--n-structures", type=int, default=250)
    ap.add_argument("--n-test", type=int, default=50)
    ap.add_argument("--grid", type=int, default=160)
    ap.add_argument("--n-init", type=int, default=32)
    ap.add_argument("--n-pick", type=int, default=24)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--out", default="search_report.json")
    a = ap.parse_args(argv)
    train, test = search_data(a.n_structures, a.n_test)
    n_tr = a.n_structures - a.n_test
    prec = capi.PLM_MIXED if a.precision == "mixed" else capi.PLM_FP64
    cs = cost_cell()
    cands = grid(a.grid)
    done = {}

    def ev(c):
        r = evaluate(c, train, test, precision=prec, cost_structures=cs)
        done[c] = r
        print(f"{c.key}: F={r.n_features} rmse_e={r.rmse_e:.3f} meV/atom "
              f"rmse_f={r.rmse_f:.4f} eV/A cost={r.cost:.5f} ms/atom ({r.seconds:.1f} s)",
              flush=True)
        return r

    rep = run_search(cands, ev, n_init=a.n_init, n_pick=a.n_pick)
    out = {
        "grid": len(cands), "valid": sum(1 for r in rep.exhaustive if r.ok),
        "data": f"{n_tr} train / {a.n_test} test binary fcc structures (128 atoms)",
        "precision": a.precision,
        "seconds_exhaustive": rep.seconds_exhaustive, "seconds_guided": rep.seconds_guided,
        "fits_exhaustive": sum(1 for r in rep.exhaustive if r.ok),
        "fits_guided": len(rep.guided), "fits_random": len(rep.random_pick),
        "front_exhaustive": rep.front_exhaustive, "front_guided": rep.front_guided,
        "front_random": rep.front_random,
        "front_recall_guided": len(set(rep.front_guided) & set(rep.front_exhaustive)) /
        max(len(rep.front_exhaustive), 1),
        "front_recall_random": len(set(rep.front_random) & set(rep.front_exhaustive)) /
        max(len(rep.front_exhaustive), 1),
        "hypervolume": {"exhaustive": rep.hv_exhaustive, "guided": rep.hv_guided,
                        "random": rep.hv_random},
        "surrogate_loss": rep.surrogate.history,
        "results": [{"key": r.cand.key, "ok": r.ok, "F": r.n_features, "rmse_e": r.rmse_e,
                     "rmse_f": r.rmse_f, "cost_ms_per_atom": r.cost, "seconds": r.seconds}
                    for r in rep.exhaustive],
    }
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1, default=float)
    print(json.dumps({k: v for k, v in out.items() if k not in ("results", "surrogate_loss")}),
          flush=True)


if __name__ == "__main__":
    main()
