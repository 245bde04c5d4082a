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
(FP32 features, split-bf16 x3 MFMA Gram, FP64 block sums) in rows/s; the
plain FP64 form is timed in the same run. Per-kernel times come from HIP
events on the library stream (plm_profile_*).

The CPU baseline times the reference's own C++ (compiled from its sources into
oracle/_ref) building X for a sample of structures with OpenMP over
structures, plus numpy X^T X, on the host cores of this machine.
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
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": rows / (t2 - t0),
        "unit": "rows/s",
        "cores": cores,
        "kind": "reference",
        "sample": f"{n_sample} structures ({rows} rows): reference C++ X build "
                  f"{t1 - t0:.2f} s + numpy X^T X {t2 - t1:.2f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-structures", type=int, default=1000)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
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
        kern = {k: capi.profile_get(k) for k in ("anlm", "pairdata", "xrows", "gram", "xty")}
        return (t1 - t0) / steps, kern

    log("timing mixed ...")
    dt_mixed, k_mixed = run(capi.PLM_MIXED, args.steps, args.warmup)
    log(f"mixed: {dt_mixed * 1e3:.1f} ms/step, {R / dt_mixed:.4g} rows/s, kernels {k_mixed}")
    log("timing fp64 ...")
    dt_fp64, k_fp64 = run(capi.PLM_FP64, args.fp64_steps, 0)
    log(f"fp64: {dt_fp64 * 1e3:.1f} ms/step, {R / dt_fp64:.4g} rows/s, kernels {k_fp64}")

    # dominant kernel of the mixed step
    ms = {k: v[0] for k, v in k_mixed.items()}
    dom = max(ms, key=ms.get)
    ms_g, n_g = k_mixed["gram"]
    algo_flops = float(R) * F * (F + 1)  # X^T X upper triangle, multiply-add = 2 flops
    if dom == "gram":
        per_launch = 3.0 * algo_flops / n_g  # three bf16 products per element
        achieved = per_launch / (ms_g / n_g * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None,
                "kernel": "gram_bf16x3_kernel"}
    else:
        ms_x, n_x = k_mixed["xrows"]
        # algorithmic bytes of one launch: X written once in FP32 plus the
        # order parameters and pair records read
        per_launch = (float(R) * F * 4.0) / n_x
        achieved = per_launch / (ms_x / n_x * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "xrows_kernel"}

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
        "dtype": "fp32-features/bf16x3-gram/fp64-accum",
        "data": "synthetic (seeded binary fcc alloy structures)",
        "config": {"workload": "binary alloy fit: 1000 x 128-atom fcc 4x4x2, rc 7, 14 Gaussians, "
                               "gtinv order 3 max_l (8,4), max_p 3, model_type 3",
                   "n_structures": len(sts), "rows": R, "features": F},
        "fp64_rows_per_s": R / dt_fp64,
        "fp64_ms_per_step": dt_fp64 * 1e3,
        "speedup_mixed_vs_fp64": dt_fp64 / dt_mixed,
        "kernel_ms_per_step_mixed": {k: v[0] / args.steps for k, v in k_mixed.items()},
        "kernel_ms_per_step_fp64": {k: v[0] / args.fp64_steps for k, v in k_fp64.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
