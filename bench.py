"""Throughput benchmark: ODE solves/sec of the 1024 x 1024 stiff mass-action
model on one MI355X, plain FP64 method (LASSIE's RKF45 / BDF1-Newton loop).

Workload (one "step" = one full solve, every sample written):
  1024 species x 1024 reactions from the seeded generator (seed 0), k
  log-uniform in [1e-6, 1e6], x0 log-uniform in [1e-6, 1], t in [0, 100] with
  200 samples, rtol 1e-6 (tol = atol + rtol*|w|), atol 1e-12, BDF step
  regrowth on, at most 8 RKF rejections in a row.
Inputs are resident in HBM (model uploaded once, outside the timed region).

Prints one JSON line.  The CPU baseline is the LSODA path (scipy odeint, the
reference's CPU method) on the same model, one thread, full solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ODE solves/sec at rtol 1e-6, chosen form vs plain FP64 on MI355X; max rel err"
FP64_PEAK_TFLOPS = 78.6
WORKLOAD = dict(S=1024, R=1024, seed=0, k_range=(1e-6, 1e6), t_end=100.0, n_samples=200)
CONTROLLER = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)


def flops_per_solve(n, st):
    """Algorithmic FP64 work: one LU per BDF step (2n^3/3) and one pair of
    triangular solves per Newton iteration (2n^2)."""
    return st["lu_factorizations"] * (2.0 * n ** 3 / 3.0) + st["newton_iters"] * 2.0 * n * n


def cpu_baseline(m, t, out_gpu):
    import numpy as np
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import lsoda
    t0 = time.perf_counter()
    y = lsoda.solve(m, t, rtol=1e-6)
    dt = time.perf_counter() - t0
    rel = np.abs(out_gpu - y) / (np.abs(y) + 1e-12)
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "one full solve of the same 1024x1024 model with the LSODA path "
                      "(scipy odeint, rtol 1e-6, atol 1e-12, 1 thread)",
            "seconds": dt}, float(np.max(rel))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (solves are "
              "independent, a multi-GPU run would shard models across processes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from lassie_amd import _lib, synth

    w = WORKLOAD
    m = synth.stiff(w["S"], w["R"], w["seed"], w["k_range"])
    t = synth.t_vector(0.0, w["t_end"], w["n_samples"])
    dm = _lib.DeviceModel(m)
    dm.solve(t, raise_on_maxiter=False, max_iters=2, **CONTROLLER)  # code objects loaded
    for _ in range(args.warmup):
        dm.solve(t, **CONTROLLER)
    kernel_ms = []
    stats = None
    out = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, _, stats, _, _ = dm.solve(t, **CONTROLLER)
        kernel_ms.append(stats["kernel_ms"])
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    k_s = float(np.mean(kernel_ms)) / 1e3
    fl = flops_per_solve(w["S"], stats)
    achieved = fl / k_s / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1024x1024 stiff mass-action model, plain FP64 RKF45/BDF1, "
                               "t=[0,100], 200 samples, rtol 1e-6",
                   "form": "plain", "species": w["S"], "reactions": w["R"], "seed": w["seed"],
                   "controller": CONTROLLER},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "solve_kernel<4,16> (persistent, whole solve)",
                     "flops_per_launch": fl, "kernel_ms": float(np.mean(kernel_ms))},
        "solver_stats": {k: stats[k] for k in ("loop_iters", "rkf_attempts", "rkf_accepts",
                                                "bdf_steps", "newton_iters", "lu_factorizations",
                                                "lu_failures", "nan_repairs")},
    }
    if not args.no_cpu:
        cb, err = cpu_baseline(m, t, out)
        line["cpu_baseline"] = cb
        line["max_rel_err_vs_lsoda"] = err
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
