#!/usr/bin/env python
"""Throughput of the fused X + X^T X hot path on one MI355X.

Workload (binary alloy): 1000 structures, each a 128-atom fcc 4x4x2 cell
(a = 3.9 A) with random A/B occupation, volume +-8 %, displacements
sigma = 0.1 A; model cutoff 7 A, 14 Gaussians (+1 constant), gtinv order 3
with max_l (8, 4), polynomial order 3, model_type 3: F = 13,020 columns and
1000 x (1 + 6 + 3*128) = 391,000 rows of X.

A step is one pass of the hot path over the whole dataset with the
structures, translations and neighbour lists already resident in HBM: order
parameters, the X rows (energy, stress, force) in column blocks, and their
accumulation into X^T X and X^T y. The headline value is the mixed form
(FP32 features; the Gram of X quantised per 4096-row chunk and column to
three 7-bit integer slices, exact on int8 MFMA with FP64 sums) in rows/s;
the plain FP64 form is timed in the same run. Per-kernel times come from HIP
events on the library stream (plm_profile_*).

The CPU baseline times the reference's own C++ (compiled from its sources into
oracle/_ref) building X for a sample of structures with OpenMP over
structures, plus numpy X^T X, on the host cores of this machine.

Further legs in the same JSON line: PolymlpEval on a 32,000-atom cell (eager
steps with the neighbour lists rebuilt, and HIP-graph replays) with the
pair-kernel roofline and the reference's CPU evaluation; the mixed-vs-FP64
fit accuracy on a sample (the mixed fit with and without its refinement
step); and the cold end-to-end fit in both precisions (device-resident
accumulation, ridge and, for the mixed form, the refinement step; and the
read-back path calc_xtx_xty + solver_ridge).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """CPU model, usable cores and NUMA nodes of this machine."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nodes = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")
                     and d[4:].isdigit()])
    except OSError:
        nodes = None
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "usable_cores": usable, "numa_nodes": nodes}


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(spec, structures, n_sample):
    """Reference C++ X build (OpenMP over structures) + numpy X^T X on a sample."""
    from oracle import ref
    from tests.helpers import write_model_yaml

    if not ref.available():
        return None
    import tempfile

    F = 13020
    with tempfile.TemporaryDirectory() as d:
        path = write_model_yaml(spec, os.path.join(d, "m.yaml"), np.zeros(F))
        R = ref.RefModel(path)
        sts = structures[:n_sample]
        t0 = time.perf_counter()
        X = R.x_batch([s.axis for s in sts], [s.positions_cartesian for s in sts],
                      [s.types for s in sts])
        t1 = time.perf_counter()
        G = X.T @ X
        t2 = time.perf_counter()
        del G
    rows = X.shape[0]
    out = {
        "value": rows / (t2 - t0),
        "unit": "rows/s",
        "cores": cpu_threads(),
        "kind": "reference",
        "sample": f"{n_sample} structures ({rows} rows): reference C++ X build (OpenMP over "
                  f"structures) {t1 - t0:.2f} s + numpy X^T X {t2 - t1:.2f} s",
    }
    out.update(host_info())
    return out


def eval_cpu_baseline(model_yaml, st, steps=1):
    """The reference's PolymlpEval (compiled from its sources into oracle/_ref):
    NeighborHalf and the evaluation with OpenMP over atoms, on the 32,000-atom
    cell; returns atoms/s and the per-phase split of one step."""
    from oracle import ref

    if not ref.available():
        return None
    R = ref.RefModel(model_yaml, with_eval=True)
    ax, pos, types = st.axis, st.positions_cartesian, st.types
    t0 = time.perf_counter()
    for _ in range(steps):
        R.eval(ax, pos, types, use_openmp=True)
    dt = (time.perf_counter() - t0) / steps
    out = {"value": st.n_total / dt, "unit": "atoms/s", "cores": cpu_threads(),
           "kind": "reference",
           "sample": f"{steps} step(s) of the {st.n_total}-atom cell (NeighborHalf + "
                     f"PolymlpEval::eval, OpenMP): {dt:.2f} s per step"}
    out.update(host_info())
    return out


FP32_PEAK_TFLOPS = 157.3


def eval_leg(steps: int, graph_steps: int, cpu: bool):
    """PolymlpEval on the 32,000-atom bcc cell (20 x 20 x 40, a = 3.30 A,
    sigma = 0.1 A) with the coefficients of a GPU FP64 ridge fit (alpha = 1e-3)
    of the 200-structure bcc dataset. One eager step = neighbour lists (cell
    list), order parameters, per-atom term contraction and the half-pair loop
    for energy, forces and stress, with the positions already resident; the
    graph steps replay the evaluation kernels over the resident lists.

    Roofline of the dominant evaluation kernel, the half-pair loop
    (pair_lane_kernel): algorithmic
    flops = 40 n_fn n_lm_half per half pair (radial functions, Y_lm and their
    gradients, and the contraction with both atoms' prod_sum), against the
    vector peak of the precision it computes in (FP64 78.6, FP32 157.3 TFLOP/s)."""
    import tempfile

    from pypolymlp_amd import capi, datasets, mlp_dev
    from pypolymlp_amd.params import save_mlp_yaml

    spec = datasets.config1_model()
    ds = datasets.config1_dataset(200)
    model = capi.Model(spec.as_dict())
    ne = mlp_dev.calc_xtx_xty(model, [ds], precision=capi.PLM_FP64)
    coef = mlp_dev.solver_ridge(ne.xtx, ne.xty, np.array([1e-3]))[:, 0]
    model.set_coeffs(coef / ne.scales)
    st = datasets.config4_structure()
    ax, pos, types = st.axis, st.positions_cartesian, st.types
    n_fn = len(spec.pair_params)
    n_lmh = (spec.max_l + 1) * (spec.max_l + 2) // 2
    out = {}
    res = {}
    for name, prec in (("fp64", capi.PLM_FP64), ("mixed", capi.PLM_MIXED)):
        res[name] = capi.eval_structure(model, ax, pos, types, prec)  # warm-up
        capi.profile_reset()
        capi.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            capi.eval_structure(model, ax, pos, types, prec)
        dt = (time.perf_counter() - t0) / steps
        capi.profile_enable(False)
        kern = {k: capi.profile_get(k) for k in ("eval_anlm", "eval_prodsum", "eval_pair")}
        out[f"eval_atoms_per_s_{name}"] = st.n_total / dt
        out[f"eval_ms_per_step_{name}"] = dt * 1e3
        out[f"eval_kernel_ms_per_step_{name}"] = {k: v[0] / steps for k, v in kern.items()}
        log(f"eval {name}: {dt * 1e3:.2f} ms/step, {st.n_total / dt:.4g} atoms/s, kernels {kern}")
        # the same evaluation as a captured HIP graph over the resident cell and
        # neighbour lists, replayed (no list rebuild between steps)
        batch = capi.Batch(model, [ax], [pos], [types])
        n_half = batch.n_pairs // 2  # one cell without self images: each pair twice
        graph = capi.EvalGraph(model, batch, prec)
        graph.run(1)
        t0 = time.perf_counter()
        graph.run(graph_steps)
        dt = (time.perf_counter() - t0) / graph_steps
        out[f"eval_graph_atoms_per_s_{name}"] = st.n_total / dt
        out[f"eval_graph_ms_per_step_{name}"] = dt * 1e3
        log(f"eval graph {name}: {dt * 1e3:.2f} ms/step over {graph_steps} replays, "
            f"{st.n_total / dt:.4g} atoms/s")
        ms_p, n_p = kern["eval_pair"]
        peak = FP64_PEAK_TFLOPS if name == "fp64" else FP32_PEAK_TFLOPS
        ach = 40.0 * n_fn * n_lmh * n_half / (ms_p / n_p * 1e-3) / 1e12
        form = os.environ.get("PLM_EVAL_PAIR") or "lane"
        out[f"eval_roofline_{name}"] = {
            "bound": "valu", "unit": "TFLOP/s",
            "kernel": "pair_lane_kernel" if form == "lane" else "pair_rec_kernel",
            "achieved": ach, "peak": peak, "frac": ach / peak,
            "work": f"40 n_fn n_lm_half = {40 * n_fn * n_lmh} flops per half pair x {n_half} "
                    "half pairs per launch"}
        del graph, batch
    e64, f64, s64 = res["fp64"]
    em, fm, sm = res["mixed"]
    out["eval_mixed_vs_fp64"] = {
        "energy_meV_per_atom": abs(em - e64) / st.n_total * 1e3,
        "force_max_meV_per_A": float(np.abs(fm - f64).max() * 1e3),
        "stress_max_rel": float(np.abs(sm - s64).max() / np.abs(s64).max()),
    }
    out["eval_workload"] = (f"{st.n_total}-atom bcc 20x20x40, a=3.30, sigma=0.1; cfg1 model "
                            "(rc 6, 11 radial, gtinv order 2, max_l 4, max_p 2) fitted on the GPU; "
                            "eval_*: neighbour lists + evaluation per step; eval_graph_*: the "
                            f"evaluation kernels captured as a HIP graph and replayed {graph_steps} "
                            "times over the resident lists")
    if cpu:
        log("eval cpu baseline ...")
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "cfg1.yaml")
            save_mlp_yaml(spec, coef / ne.scales, np.ones(len(coef)), path)
            out["eval_cpu_baseline"] = eval_cpu_baseline(path, st)
        log(f"eval cpu: {out['eval_cpu_baseline']}")
    return out


def fit_accuracy_leg(n_st: int):
    """Mixed against FP64 fit on a seeded subset of the same binary-alloy
    workload (with Morse targets): the device-resident normal equations of
    both forms, the GPU ridge solve at alpha = 1e-3 and 1e-1, for the mixed
    form with and without the refinement step that mlp_dev.fit applies, and
    FP64 evaluation of the held-out 10 %. Returns the largest energy
    (meV/atom) and force (meV/A) differences."""
    from pypolymlp_amd import capi, datasets, mlp_dev

    ds = datasets.config2_dataset(n_st)
    train, test = ds.split(0.9)
    m = capi.Model(datasets.config2_model().as_dict())
    alphas = np.array([1e-3, 1e-1])
    fits = {}
    for name, prec in (("fp64", capi.PLM_FP64), ("mixed", capi.PLM_MIXED)):
        dn = mlp_dev.accumulate_normal(m, [train], precision=prec, keep_parts=name == "mixed")
        c, info, _ = dn.acc.ridge(dn.scales, dn.zero, alphas)
        fits[name] = c / dn.scales[:, None]
        if name == "mixed":
            c2, _, _ = mlp_dev.refine(dn, c, alphas, prec)
            fits["mixed_refined"] = c2 / dn.scales[:, None]
        del dn
    sts = test.structures
    b = capi.Batch(m, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                   [s.types for s in sts])
    n = np.array([s.n_total for s in sts])
    err = {"mixed": [0.0, 0.0], "mixed_refined": [0.0, 0.0]}
    for k in range(len(alphas)):
        pred = {}
        for name, c in fits.items():
            m.set_coeffs(c[:, k])
            E, F, _ = capi.eval_batch(m, b, capi.PLM_FP64)
            pred[name] = (E / n, np.concatenate([f.reshape(-1) for f in F]))
        for name in err:
            err[name][0] = max(err[name][0], float(np.abs(pred[name][0] - pred["fp64"][0]).max()))
            err[name][1] = max(err[name][1], float(np.abs(pred[name][1] - pred["fp64"][1]).max()))
    de, df = err["mixed_refined"]
    log(f"fit accuracy: mixed (refined) vs FP64 max |dE| {de * 1e3:.2e} meV/atom, max |dF| "
        f"{df * 1e3:.2e} meV/A; without the refinement step {err['mixed'][0] * 1e3:.2e} / "
        f"{err['mixed'][1] * 1e3:.2e}")
    return {"fit_mixed_vs_fp64": {
        "energy_max_meV_per_atom": de * 1e3, "force_max_meV_per_A": df * 1e3,
        "unrefined": {"energy_max_meV_per_atom": err["mixed"][0] * 1e3,
                      "force_max_meV_per_A": err["mixed"][1] * 1e3},
        "sample": f"{n_st} structures of the workload with Morse targets, 90/10 split, "
                  "alpha 1e-3 and 1e-1, held-out predictions; the mixed fit refined once "
                  "(mlp_dev.refine, as mlp_dev.fit does)"}}


def fit_end_to_end_leg(n_st: int, alphas=(1e-3,)):
    """The fit as a user calls it, cold: mlp_dev.calc_xtx_xty (row weights,
    batch upload with the host lattice translations and the GPU neighbour
    lists, the accumulation, read-back and scaling of the normal equations;
    data_sequential.py:25-94) followed by mlp_dev.solver_ridge (solvers.py:48-84),
    on the workload's structures, for both precisions. The targets do not
    change the work, so zero targets are used. Returns rows/s of the whole
    call and the split."""
    from pypolymlp_amd import capi, datasets, mlp_dev

    ds = datasets.config2_dataset(n_st, targets=False)
    model = capi.Model(datasets.config2_model().as_dict())
    # one-time library work (code objects, model tables) outside the timing
    mlp_dev.calc_xtx_xty(model, [ds.subset(range(2))], precision=capi.PLM_MIXED)
    rows = mlp_dev.reference_n_rows(ds)
    out = {}
    for name, prec in (("mixed", capi.PLM_MIXED), ("fp64", capi.PLM_FP64)):
        # the reference-API path: X^T X read back, scaled on the host, solved
        t0 = time.perf_counter()
        ne = mlp_dev.calc_xtx_xty(model, [ds], precision=prec)
        t1 = time.perf_counter()
        mlp_dev.solver_ridge(ne.xtx, ne.xty, np.asarray(alphas))
        t2 = time.perf_counter()
        del ne
        # the path of mlp_dev.fit: the normal equations stay on the device;
        # the quantised form adds its refinement step (a residual pass over
        # the rows and the correction solve)
        t3 = time.perf_counter()
        dn = mlp_dev.accumulate_normal(model, [ds], precision=prec, keep_parts=name == "mixed")
        t4 = time.perf_counter()
        capi.profile_reset()
        capi.profile_enable(True)
        c, _, _ = dn.acc.ridge(dn.scales, dn.zero, np.asarray(alphas))
        capi.profile_enable(False)
        t5 = time.perf_counter()
        chol_ms, n_chol = capi.profile_get("cholesky")
        if name == "mixed":
            mlp_dev.refine(dn, c, np.asarray(alphas), prec)
        t6 = time.perf_counter()
        del dn
        # Cholesky roofline: F^3 / 3 flops per factorization against the FP64
        # MFMA peak (the trailing update is FP64 MFMA work)
        Fm = model.n_features
        chol_tflops = n_chol * Fm ** 3 / 3.0 / (chol_ms * 1e-3) / 1e12 if chol_ms > 0 else None
        out[name] = {"rows_per_s": rows / (t6 - t3), "seconds": t6 - t3,
                     "accumulate_s": t4 - t3, "ridge_s": t5 - t4, "refine_s": t6 - t5,
                     "cholesky": {"ms_per_factorization": chol_ms / max(n_chol, 1),
                                  "achieved_tflops": chol_tflops, "peak_tflops": FP64_PEAK_TFLOPS,
                                  "frac": chol_tflops / FP64_PEAK_TFLOPS if chol_tflops else None,
                                  "work": f"F^3/3 flops per factorization, F = {Fm}"},
                     "read_back_path": {"rows_per_s": rows / (t2 - t0), "seconds": t2 - t0,
                                        "calc_xtx_xty_s": t1 - t0, "solver_ridge_s": t2 - t1}}
        log(f"end-to-end {name}: device-resident {t6 - t3:.2f} s ({rows / (t6 - t3):.4g} rows/s;"
            f" accumulate {t4 - t3:.2f} s, ridge {t5 - t4:.2f} s, refine {t6 - t5:.2f} s); "
            f"read-back path {t2 - t0:.2f} s (calc_xtx_xty {t1 - t0:.2f} s, ridge {t2 - t1:.2f} s)")
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["workload"] = (f"{n_st} structures of the workload, cold: row weights, batch upload and "
                       "neighbour lists, accumulation, scaling and the ridge solve at alpha "
                       f"{list(alphas)} (mlp_dev.accumulate_normal + plm_normal_ridge, the path "
                       "of mlp_dev.fit), plus for the mixed form its refinement step "
                       "(mlp_dev.refine); read_back_path: calc_xtx_xty + solver_ridge")
    return {"fit_end_to_end": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-structures", type=int, default=1000)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--graph-steps", type=int, default=2,
                    help="replays of the mixed step captured as a HIP graph (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--eval-steps", type=int, default=5)
    ap.add_argument("--eval-graph-steps", type=int, default=1000)
    ap.add_argument("--e2e-structures", type=int, default=1000,
                    help="structures of the cold end-to-end fit timing (0: skip)")
    ap.add_argument("--fit-sample", type=int, default=120,
                    help="structures of the mixed-vs-FP64 fit accuracy check (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three plain 1 GiB device copies (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        x = torch.ones(256 * 1024 * 1024, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 3 x 1 GiB")
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path here runs on a single MI355X; "
              "multi-GPU scaling is not part of this build.", file=sys.stderr)
        sys.exit(2)

    from pypolymlp_amd import capi, datasets

    spec = datasets.config2_model()
    t = time.perf_counter()
    ds = datasets.config2_dataset(args.n_structures, targets=False)
    log(f"dataset: {len(ds.structures)} structures in {time.perf_counter() - t:.1f} s")
    model = capi.Model(spec.as_dict())
    F = model.n_features
    sts = ds.structures
    t = time.perf_counter()
    batch = capi.Batch(model, [s.axis for s in sts], [s.positions_cartesian for s in sts],
                       [s.types for s in sts])
    R = batch.n_rows
    log(f"batch resident: {R} rows, F = {F}, setup {time.perf_counter() - t:.1f} s")
    w = np.ones(R)
    y = np.zeros(R)

    def run(prec, steps, warmup):
        acc = capi.Normal(model)
        for _ in range(warmup):
            acc.accumulate(batch, w, y, prec)
        capi.profile_reset()
        capi.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            acc.accumulate(batch, w, y, prec)
        t1 = time.perf_counter()
        capi.profile_enable(False)
        kern = {k: capi.profile_get(k) for k in ("anlm", "wprime", "pairdata", "xrows", "xgroup",
                                                 "xenergy", "poly_deriv", "poly_z", "poly_rows",
                                                 "quantize", "gram")}
        return (t1 - t0) / steps, kern

    log("timing mixed ...")
    dt_mixed, k_mixed = run(capi.PLM_MIXED, args.steps, args.warmup)
    log(f"mixed: {dt_mixed * 1e3:.1f} ms/step, {R / dt_mixed:.4g} rows/s, kernels {k_mixed}")
    # the same step captured once as a HIP graph and replayed (plm_normal_graph_*)
    dt_graph = None
    if args.graph_steps > 0:
        acc_g = capi.Normal(model)
        graph = capi.NormalGraph(acc_g, batch, w, y, capi.PLM_MIXED)
        graph.run(1)
        t0 = time.perf_counter()
        graph.run(args.graph_steps)
        dt_graph = (time.perf_counter() - t0) / args.graph_steps
        del graph, acc_g
        log(f"mixed as a HIP graph: {dt_graph * 1e3:.1f} ms/step over {args.graph_steps} replays")
    dt_fp64, k_fp64 = None, {}
    if args.fp64_steps > 0:
        log("timing fp64 ...")
        # one untimed step first: the autotuner times its launch shapes on the
        # first pass of each precision
        dt_fp64, k_fp64 = run(capi.PLM_FP64, args.fp64_steps, 1)
        log(f"fp64: {dt_fp64 * 1e3:.1f} ms/step, {R / dt_fp64:.4g} rows/s, kernels {k_fp64}")

    # Rooflines. The headline kernel is the Gram of [X | y] (F' = F + 1
    # columns): algorithmic work F'(F'+1) ops per row (the upper triangle,
    # 2 ops per product). The mixed form computes the exact Gram of X quantised
    # to three 7-bit slices, which takes 9 int8 slice products per algorithmic
    # product, so its method peak is the int8 MFMA peak / 9 (the convention of
    # an n-product split Gram: algorithmic work against peak / n); the issued
    # int8 rate is reported beside it. The FP64 form runs against the FP64
    # MFMA peak. Profile totals cover all timed steps.
    Fy = F + 1
    alg_per_row = float(Fy) * (Fy + 1)
    ms_g, n_g = k_mixed["gram"]
    ach = alg_per_row * R * args.steps / (ms_g * 1e-3) / 1e12
    gram_i8 = {"bound": "mfma", "kernel": "gram_i8_kernel",
               "unit": "TOP/s (algorithmic Gram ops of the exact 3-slice int8 method)",
               "achieved": ach, "peak": INT8_PEAK_TOPS / 9.0, "frac": ach / (INT8_PEAK_TOPS / 9.0),
               "launches_per_step": n_g / args.steps,
               "work": f"F'(F'+1) = {alg_per_row:.4g} ops per row, F' = {Fy}; {R} rows per step",
               "issued_int8": {"achieved_tops": 9.0 * ach, "peak_tops": INT8_PEAK_TOPS,
                               "frac": 9.0 * ach / INT8_PEAK_TOPS},
               "vs_fp64_mfma_peak": ach / FP64_PEAK_TFLOPS}
    gram_f64 = None
    if args.fp64_steps > 0:
        ms_g64, n_g64 = k_fp64["gram"]
        a64 = alg_per_row * R * args.fp64_steps / (ms_g64 * 1e-3) / 1e12
        gram_f64 = {"bound": "mfma", "kernel": "gram_f64_kernel", "unit": "TFLOP/s (fp64)",
                    "achieved": a64, "peak": FP64_PEAK_TFLOPS, "frac": a64 / FP64_PEAK_TFLOPS,
                    "launches_per_step": n_g64 / args.fp64_steps}
    # linear columns of X by signature groups (xgroup_kernel on the static
    # grid, xgroup_queue_kernel on the persistent one, as the autotuner chose):
    # algorithmic bytes = its output (stress and force rows of the linear
    # columns, FP32) plus its inputs read once (the pair records of the step)
    n_fn = len(spec.pair_params)
    n_lmh = (spec.max_l + 1) * (spec.max_l + 2) // 2
    rec_bytes = 4.0 * ((8 + 2 * n_fn + 8 * n_lmh + 3) // 4 * 4)
    n_lin = len(capi.feature_attrs(model)[0])
    xg_bytes = float(R - len(sts)) * n_lin * 4.0 + batch.n_pairs * rec_bytes
    ms_x, n_x = k_mixed["xgroup"]
    tuned = [ln for ln in capi.tuning_report(model).splitlines() if "fp32" in ln]
    forced = os.environ.get("PLM_XGROUP", "")
    persistent = forced.endswith(",1") if forced else bool(tuned) and "persistent" in tuned[-1]
    xg_kernel = "xgroup_queue_kernel" if persistent else "xgroup_kernel"
    xgroup = {"bound": "hbm", "kernel": xg_kernel, "unit": "GB/s", "peak": HBM_PEAK_GBS,
              "launches_per_step": n_x / args.steps,
              "achieved": xg_bytes * args.steps / (ms_x * 1e-3) / 1e9 if ms_x else None}
    if xgroup["achieved"]:
        xgroup["frac"] = xgroup["achieved"] / HBM_PEAK_GBS
    roof = dict(gram_i8)
    # HBM bytes per launch of the Gram: rocprofv3 FETCH_SIZE and WRITE_SIZE of
    # this workload (scaled by a 1 GiB copy), committed with the profiles; the
    # counters cannot be read from inside this process
    roof["traffic"] = None
    tpath = os.path.join(ROOT, "profiles", "fit_binary_alloy_traffic.json")
    if os.path.exists(tpath) and len(sts) == 1000:
        with open(tpath) as f:
            tr = json.load(f)
        k = tr.get(roof["kernel"])
        if k:
            roof["traffic"] = k["fetch_bytes_per_launch"] + k["write_bytes_per_launch"]
            roof["traffic_source"] = "profiles/fit_binary_alloy_traffic.json (rocprofv3 --pmc)"
            alg_bytes = 3.0 * Fy * R / max(n_g / args.steps, 1)  # three int8 planes per launch
            roof["traffic_vs_algorithmic"] = roof["traffic"] / alg_bytes

    ev = eval_leg(args.eval_steps, args.eval_graph_steps, not args.no_cpu_baseline) \
        if args.eval_steps > 0 else {}
    if args.fit_sample > 0:
        ev.update(fit_accuracy_leg(args.fit_sample))
    if args.e2e_structures > 0:
        ev.update(fit_end_to_end_leg(args.e2e_structures))

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(spec, sts, args.cpu_sample)
        if cpu:
            log(f"cpu: {cpu}")

    out = {
        "metric": "fused X+XtX rows/s, mixed vs FP64 at <=0.1 meV/atom; eval atoms/s",
        "value": R / dt_mixed,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 features; int8x3-slice exact Gram, fp64 sums",
        "data": "synthetic (seeded binary fcc alloy structures)",
        "config": {"workload": "binary alloy fit: 1000 x 128-atom fcc 4x4x2, rc 7, 14 Gaussians, "
                               "gtinv order 3 max_l (8,4), max_p 3, model_type 3",
                   "n_structures": len(sts), "rows": R, "features": F},
        "fp64_rows_per_s": R / dt_fp64 if dt_fp64 else None,
        "fp64_ms_per_step": dt_fp64 * 1e3 if dt_fp64 else None,
        "speedup_mixed_vs_fp64": dt_fp64 / dt_mixed if dt_fp64 else None,
        "graph_ms_per_step_mixed": dt_graph * 1e3 if dt_graph else None,
        "kernel_ms_per_step_mixed": {k: v[0] / args.steps for k, v in k_mixed.items()},
        "kernel_ms_per_step_fp64": {k: v[0] / args.fp64_steps for k, v in k_fp64.items()},
        "roofline": roof,
        "roofline_kernels": {"gram_i8": gram_i8, "gram_f64": gram_f64, "xgroup": xgroup},
        "xgroup_tuning": capi.tuning_report(model).strip().splitlines(),
        "cpu_baseline": cpu,
    }
    out.update(ev)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
