"""Throughput benchmark: ODE solves/sec on one MI355X, LASSIE's plain FP64
method (RKF45 / BDF1-Newton loop), plus the chosen (batched) form.

Headline workload (one "step" = one full solve, every sample written):
  1024 species x 1024 reactions from the seeded generator (seed 0), k
  log-uniform in [1e-6, 1e6], x0 log-uniform in [1e-6, 1], t in [0, 100] with
  200 samples, rtol 1e-6 (tol = atol + rtol*|w|), atol 1e-12, BDF step
  regrowth on, at most 8 RKF rejections in a row.  The model is uploaded once,
  outside the timed region; the solve runs as one persistent kernel.

Chosen form (reported beside the headline, not as `value`): the batched
persistent kernel, prepared as a plan (graph replay, autotuned kernel variant),
on uneven instances (S in {16..512}, R = 2S, k logU[1e-4, 1e4], t in [0, 50],
100 samples), against the plain single-model path on a subsample of the same
instances (their trajectories must agree bit for bit).

Differentiation form (also beside the headline): forward pass with step record
plus discrete-adjoint gradient of a log-loss, for B parameterisations of one
512 x 1024 network on a short horizon.

CPU baselines: the LSODA path (scipy odeint, the reference's CPU method) on the
same 1024x1024 model, one thread, full solve; and the oracle (the same
algorithm on the CPU, one thread) projected from its measured per-operation
cost and the GPU run's operation counts.
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
    """Algorithmic FP64 work of one solve: one LU per BDF step (2n^3/3) and one
    pair of triangular solves per Newton iteration (2n^2)."""
    return st["lu_factorizations"] * (2.0 * n ** 3 / 3.0) + st["newton_iters"] * 2.0 * n * n


def cpu_lsoda(m, t, out_gpu):
    import numpy as np
    from oracle import lsoda
    t0 = time.perf_counter()
    y = lsoda.solve(m, t, rtol=1e-6)
    dt = time.perf_counter() - t0
    rel = np.abs(out_gpu - y) / (np.abs(y) + 1e-12)
    sig = np.abs(y) > 1e-6 * np.abs(y).max()
    return dt, float(rel.max()), float(np.median(rel[sig]))


def cpu_same_algorithm(m, st):
    """Per-operation cost of the oracle on this model (one LU + solve at the
    initial state, one RKF45 attempt), scaled by the GPU run's counts."""
    import numpy as np
    from oracle import oracle as O
    x = np.asarray(m.x0, float)
    t0 = time.perf_counter()
    O.rkf_attempt(m, x, 1e-3, rtol=1e-6)
    t_rkf = time.perf_counter() - t0
    J = O.jac(m, x)
    t0 = time.perf_counter()
    M = O.newton_matrix(O.jac(m, x), 0.01)
    LU, piv, _ = O.lu(M)
    t_lu = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(3):
        O.lu_solve(LU, piv, x)
        O.rhs(m, x)
    t_sol = (time.perf_counter() - t0) / 3
    del J
    proj = st["bdf_steps"] * t_lu + st["newton_iters"] * t_sol + st["rkf_attempts"] * t_rkf
    return proj, dict(lu_s=t_lu, newton_iter_s=t_sol, rkf_attempt_s=t_rkf)


def cpu_batched(models, t, st, sample):
    """CPU baselines of the batched form, on one core (this process has
    initialised the GPU, so it starts no worker processes): LSODA (the
    reference's CPU method, scipy odeint) on the first `sample` instances, and
    the oracle's per-operation cost for each size times the GPU run's
    per-instance operation counts (the same algorithm)."""
    import numpy as np
    from oracle import lsoda
    from oracle import oracle as O
    t0 = time.perf_counter()
    for m in models[:sample]:
        lsoda.solve(m, t, rtol=1e-6)
    wall = time.perf_counter() - t0
    unit = {}
    for m in models:
        if m.S in unit:
            continue
        x = np.asarray(m.x0, float)
        t0 = time.perf_counter()
        O.rkf_attempt(m, x, 1e-3, rtol=1e-6)
        t_rkf = time.perf_counter() - t0
        t0 = time.perf_counter()
        LU, piv, _ = O.lu(O.newton_matrix(O.jac(m, x), 0.01))
        t_lu = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(3):
            O.lu_solve(LU, piv, x)
            O.rhs(m, x)
        unit[m.S] = (t_lu, (time.perf_counter() - t0) / 3, t_rkf)
    proj = sum(d["bdf_steps"] * unit[m.S][0] + d["newton_iters"] * unit[m.S][1]
               + d["rkf_attempts"] * unit[m.S][2] for m, d in zip(models, st))
    return {"lsoda": {"value": sample / wall, "unit": "solves/s", "cores": 1, "kind": "port",
                      "sample": f"the first {sample} of the {len(models)} instances, LSODA (scipy odeint "
                                f"restating CPU-based LSODA/LSODA.py; rtol 1e-6, atol 1e-12), one core; "
                                f"independent instances, so N cores give N times this",
                      "seconds": wall},
            "same_algorithm_projection": {
                "value": len(models) / proj, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "oracle per-operation cost per size (one LU, one Newton solve + RHS, one "
                          "RKF45 attempt) x the GPU run's per-instance counts", "projected_seconds": proj}}


def chosen_form(n_batch, n_plain, with_plain, with_cpu=False):
    """Batched FP64 form (a plan: graph replay, autotuned kernel variant) vs
    the plain path on the same instances."""
    from lassie_amd import _lib, synth
    t0 = time.perf_counter()
    models = synth.uneven_batch(n_batch, 1000)
    t = synth.t_vector(0.0, 50.0, 100)
    say(f"  {n_batch} models generated ({time.perf_counter() - t0:.1f} s)")
    dms = []
    for i, m in enumerate(models):
        dms.append(_lib.DeviceModel(m))
        if (i + 1) % 256 == 0:
            say(f"  {i + 1} models uploaded ({time.perf_counter() - t0:.1f} s)")
    plan = _lib.BatchPlan(dms, t, **CONTROLLER)
    variant, tile_rows, trial_ms = plan.info()
    say(f"  plan prepared, variant {variant} ({time.perf_counter() - t0:.1f} s)")
    # the run blocks this thread; another one reports each time the device's
    # count of finished instances has grown by 1/64 of the batch
    import threading
    finished = threading.Event()

    def follow(t_start):
        step, last = max(1, n_batch // 64), 0
        while not finished.wait(2.0):
            done = plan.progress()
            if done >= last + step:
                last = done
                say(f"  {done} of {n_batch} instances finished ({time.perf_counter() - t_start:.1f} s)")

    t0 = time.perf_counter()
    watcher = threading.Thread(target=follow, args=(t0,), daemon=True)
    watcher.start()
    try:
        outs, _, st, ms, rc = plan.run()
    finally:
        finished.set()
        watcher.join()
    wall = time.perf_counter() - t0
    say(f"  batched run: {wall:.2f} s")
    plan.close()
    res = {"instances": n_batch, "sizes": sorted({m.S for m in models}),
           "batched_solves_per_s": n_batch / wall, "batched_device_ms": ms,
           "precision": "f64 (bit-identical to the oracle)",
           "autotuner": {"variant": variant, "lu_tile_rows_in_flight": tile_rows,
                         "trial_ms_per_variant": trial_ms},
           "launch": "hipGraph replay of a prepared plan",
           "note": "C4 names 2048 instances; --batch 2048 measures them, but its makespan is set by a "
                   "few instances that the LASSIE controller drives into tens of millions of "
                   "modified-Newton iterations, so the default measures the first 256"}
    if with_plain:
        plain_s = 0.0
        for dm, ob in zip(dms[:n_plain], outs[:n_plain]):
            t0 = time.perf_counter()
            out, _, _, _, _ = dm.solve(t, **CONTROLLER)
            plain_s += time.perf_counter() - t0
            assert (out == ob).all(), "batched and plain forms disagree"
        res["plain_subsample"] = n_plain
        res["plain_solves_per_s"] = n_plain / plain_s
        res["speedup_vs_plain"] = res["batched_solves_per_s"] / res["plain_solves_per_s"]
        res["max_rel_err_vs_plain"] = 0.0
    for dm in dms:
        dm.close()
    if with_cpu:
        say("  CPU baselines of the batched form (one core)")
        res["cpu_baseline"] = cpu_batched(models, t, st, min(32, n_batch))
    return res


def differentiation_form(B, t_end):
    """Forward + discrete-adjoint gradient of the log-loss for B
    parameterisations of one 512 x 1024 network (k logU[1e-4, 1e4], 50
    samples on [0, t_end]); reports gradients/s."""
    import numpy as np
    from lassie_amd import _lib, synth
    models, hidden = synth.parameterisations(512, 1024, B, seed=0)
    t = synth.t_vector(0.0, t_end, 50)
    dms = [_lib.DeviceModel(m) for m in models]
    eps = 1e-12
    t0 = time.perf_counter()
    outs, st, fwd_ms, tape = _lib.forward_batch(dms, t, max_steps=1 << 16, **CONTROLLER)
    target = [o * np.exp(0.1) for o in outs]  # fixed synthetic target
    gouts = [2.0 * (np.log(o + eps) - np.log(tg + eps)) / (o + eps) for o, tg in zip(outs, target)]
    gks, _, adj_ms = tape.adjoint(gouts)
    wall = time.perf_counter() - t0
    steps = [tape.info(i)[0] for i in range(B)]
    tape.close()
    for dm in dms:
        dm.close()
    return {"parameterisations": B, "network": "512 species x 1024 reactions", "t": [0.0, t_end],
            "samples": 50, "gradients_per_s": B / wall, "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
            "mean_steps": float(np.mean(steps)), "checkpoint_every": 32,
            "finite_fraction": float(np.mean([np.isfinite(g).all() for g in gks]))}


def precision_form_4096(cap):
    """The 4096 x 4096 configuration (same generator as the headline, seed 0)
    in the plain FP64 form and in the single-model precision form (FP32 grid
    LU, FP64 refinement to a componentwise 1e-14 with the species' atol as
    floor), over the first `cap` loop iterations of the same controller."""
    import numpy as np
    from lassie_amd import _lib, synth
    m = synth.stiff(4096, 4096, 0, WORKLOAD["k_range"])
    t = synth.t_vector(0.0, WORKLOAD["t_end"], WORKLOAD["n_samples"])
    dm = _lib.DeviceModel(m)
    res, finals = {}, {}
    for name, kw in (("fp64", {}), ("fp32_refine", dict(precision=_lib.PREC_FP32_REFINE))):
        say(f"  4096x4096 {name}, first {cap} loop iterations")
        t0 = time.perf_counter()
        _, yf, st, rc, _ = dm.solve(t, raise_on_maxiter=False, max_iters=cap, **CONTROLLER, **kw)
        wall = time.perf_counter() - t0
        finals[name] = yf
        lu = max(st["lu_factorizations"], 1)
        res[name] = {"wall_s": wall, "kernel_ms": st["kernel_ms"], "bdf_steps": st["bdf_steps"],
                     "newton_iters": st["newton_iters"], "lu_factorizations": st["lu_factorizations"],
                     "lu_ms_per_factorization": st["phase_ms"][2] / lu,
                     "newton_ms_per_iteration": st["phase_ms"][3] / max(st["newton_iters"], 1),
                     "refine_sweeps": st["refine_iters"], "refine_fallbacks": st["refine_fallbacks"],
                     "groups": st["groups"]}
        say(f"  {name}: {wall:.1f} s")
    dm.close()
    a, b = finals["fp64"], finals["fp32_refine"]
    res["max_rel_err_final_state"] = float(np.max(np.abs(b - a) / (np.abs(a) + m.atol)))
    res["loop_iterations"] = cap
    return res


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-chosen", action="store_true", help="skip the batched-form comparison")
    ap.add_argument("--batch", type=int, default=256, help="instances in the batched-form run")
    ap.add_argument("--plain-subsample", type=int, default=8)
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation-form measurement")
    ap.add_argument("--no-fp32", action="store_true", help="skip the single-precision loop (-p float)")
    ap.add_argument("--diff-batch", type=int, default=256)
    ap.add_argument("--diff-t-end", type=float, default=0.2)
    ap.add_argument("--precision-4096", type=int, default=0, metavar="N",
                    help="also compare the plain and precision forms on the 4096x4096 model over "
                         "its first N loop iterations (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: solves are "
              "independent, so several GPUs would each run their own share of models in separate "
              "processes, with no collective in the solve path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from lassie_amd import _lib, synth

    w = WORKLOAD
    m = synth.stiff(w["S"], w["R"], w["seed"], w["k_range"])
    t = synth.t_vector(0.0, w["t_end"], w["n_samples"])
    say(f"plain FP64 solve, {w['S']}x{w['R']}, {args.steps} step(s) after {args.warmup} warmup")
    dm = _lib.DeviceModel(m)
    for _ in range(args.warmup):
        dm.solve(t, **CONTROLLER)
    kernel_ms, stats, out = [], None, None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, _, stats, _, _ = dm.solve(t, **CONTROLLER)
        kernel_ms.append(stats["kernel_ms"])
    elapsed = time.perf_counter() - t0
    dm.close()
    k_ms = float(np.mean(kernel_ms))
    fl = flops_per_solve(w["S"], stats)
    achieved = fl / (k_ms / 1e3) / 1e12
    line = {
        "metric": METRIC, "value": args.steps / elapsed, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1024x1024 stiff mass-action model, plain FP64 RKF45/BDF1, "
                               "t=[0,100], 200 samples, rtol 1e-6",
                   "form": "plain", "species": w["S"], "reactions": w["R"], "seed": w["seed"],
                   "k_range": list(w["k_range"]), "controller": CONTROLLER},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "solve_kernel<4,16,double,false> (persistent: one launch = one solve)",
                     "flops_per_launch": fl, "kernel_ms": k_ms,
                     "phase_ms": {"control_rkf": stats["phase_ms"][0],
                                  "jacobian": stats["phase_ms"][1], "lu": stats["phase_ms"][2],
                                  "newton_solves": stats["phase_ms"][3]}},
        "solver_stats": {k: stats[k] for k in ("loop_iters", "rkf_attempts", "rkf_accepts",
                                                "bdf_steps", "newton_iters", "lu_factorizations",
                                                "lu_failures", "nan_repairs", "groups")},
    }
    traffic = os.path.join(ROOT, "profiles", "traffic_solve_1024.json")
    if os.path.exists(traffic):
        with open(traffic) as f:
            line["roofline"]["traffic"] = json.load(f).get("bytes_per_launch")
    say(f"plain solve done: {elapsed:.2f} s")
    if not args.no_cpu:
        say("CPU LSODA baseline")
        t_lsoda, err_max, err_med = cpu_lsoda(m, t, out)
        proj, unit = cpu_same_algorithm(m, stats)
        line["cpu_baseline"] = {
            "value": 1.0 / t_lsoda, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "one full solve of the same 1024x1024 model by the LSODA path (scipy odeint "
                      "restating CPU-based LSODA/LSODAsim.py; rtol 1e-6, atol 1e-12, 1 thread)",
            "seconds": t_lsoda}
        line["cpu_same_algorithm"] = {
            "value": 1.0 / proj, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle (CPU restatement of the same RKF45/BDF1 loop) per-operation cost on "
                      "this model x the GPU run's operation counts", "projected_seconds": proj,
            **unit}
        line["max_rel_err_vs_lsoda"] = err_max
        line["median_rel_err_vs_lsoda"] = err_med
    if not args.no_fp32:
        say("single-precision loop (-p float) on the same model")
        dm = _lib.DeviceModel(m)
        t0 = time.perf_counter()
        out32, _, st32, rc32, _ = dm.solve(t, precision=_lib.PREC_FP32_LOOP, **CONTROLLER)
        el32 = time.perf_counter() - t0
        dm.close()
        rel = np.abs(out32 - out) / (np.abs(out) + m.atol[None, :])
        line["fp32_loop"] = {
            "solves_per_s": 1.0 / el32, "kernel_ms": st32["kernel_ms"],
            "lu_ms_per_factorization": st32["phase_ms"][2] / max(st32["lu_factorizations"], 1),
            "bdf_steps": st32["bdf_steps"], "newton_iters": st32["newton_iters"],
            "max_rel_err_vs_fp64": float(rel.max()), "median_rel_err_vs_fp64": float(np.median(rel)),
            "note": "Simulation<float> semantics (the reference's -p float), bit-identical to the "
                    "single-precision oracle"}
        say(f"single-precision loop done: {el32:.2f} s")
    if not args.no_chosen:
        say(f"batched form, {args.batch} uneven instances; plain path on {args.plain_subsample}")
        line["chosen_form"] = chosen_form(args.batch, args.plain_subsample, True, not args.no_cpu)
    if not args.no_diff:
        say(f"differentiation form, {args.diff_batch} parameterisations of 512x1024, t in [0, {args.diff_t_end}]")
        line["differentiation_form"] = differentiation_form(args.diff_batch, args.diff_t_end)
    if args.precision_4096 > 0:
        say(f"precision form on 4096x4096, {args.precision_4096} loop iterations")
        line["precision_form_4096"] = precision_form_4096(args.precision_4096)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
