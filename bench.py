"""Throughput benchmark: ODE solves/sec on one MI355X, the chosen form (the
batched FP64 plan) against LASSIE's plain FP64 method on the same models.

Workload (one "step" = one pass over the whole batch, every sample written):
  2048 independent instances of the uneven stiff generator, S drawn uniformly
  from {16, 32, 64, 128, 256, 512}, R = 2S, instance i from seed 1000 + i,
  k log-uniform in [1e-4, 1e4], x0 log-uniform in [1e-6, 1]; t in [0, 50] with
  100 samples; controller rtol 1e-6 (tol = atol + rtol*|w|), atol 1e-12, BDF
  step regrowth, at most 8 RKF rejections in a row, and at most 50000
  modified-Newton iterations per instance (the same cap in every form: about
  1% of the instances reach it, see `capped`).  Models are uploaded and the
  plan prepared outside the timed region; a step replays the plan (the
  workgroup-level kernel for S > 64 and, concurrently, the wave-level kernel
  for S <= 64) and copies every trajectory back.

Plain form: the single-model persistent solver (one launch per instance), run
on a stratified sample of the same instances (the first instances of every
size); its trajectories must equal the chosen form's bit for bit, and its
workload rate is projected from the per-size mean times.

Roofline: the workgroup-level kernel (the dominant one): algorithmic FP64
work (2S^3/3 per LU, 2S^2 per Newton solve) of its instances over its
device time from HIP events on its stream, against the FP64 peak; `traffic`
is its measured HBM bytes per launch (profiles/traffic_batch_c4.json).

Differentiation leg (the discrete adjoint, one line field): 256
parameterisations of one 512-species x 1024-reaction stiff network (k
logU[1e-4, 1e4], one seed each), t in [0, 0.2] with 50 samples; the target
trajectories come from the hidden rates k * U[0.5, 2] (hidden seed 7) solved
by the batched form; one gradient evaluation = the forward pass with its step
record + the log-loss cotangent + the reverse pass, for all 256 at once.
Agreement of these gradients with central differences (h = 1e-6) and with
the oracle's adjoint is checked by the GPU tests.

CPU baselines (rank 0 only, before the GPU is touched, one core each):
  cpu_baseline   the oracle (the same FP64 RKF45/BDF1 algorithm on the CPU) on a
                 bounded stratified sample: full solves for S <= 128, a prefix
                 of the loop for S = 256 and 512 scaled by each instance's full
                 BDF-step count from the GPU run;
  cpu_lsoda      LSODA (scipy odeint, restating the reference's CPU-based
                 LSODA/LSODAsim.py) on the same sample, a different algorithm.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ODE solves/sec at rtol 1e-6, chosen form vs plain FP64 on MI355X; max rel err"
FP64_PEAK_TFLOPS = 78.6
WORKLOAD = dict(n=2048, seed_base=1000, sizes=(16, 32, 64, 128, 256, 512), k_range=(1e-4, 1e4),
                t_end=50.0, n_samples=100)
CONTROLLER = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8, max_newton=50000)
DIFF = dict(S=512, R=1024, B=256, seed=0, t_end=0.2, n_samples=50, eps=1e-12)
DIFF_CTRL = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
T_START = time.perf_counter()


def say(msg):
    print(f"[bench {time.perf_counter() - T_START:6.1f}s] {msg}", file=sys.stderr, flush=True)


def elapsed():
    return time.perf_counter() - T_START


def work_flops(S, st):
    """Algorithmic FP64 work of one instance: one LU per factorization
    (2S^3/3) and one pair of triangular solves per Newton iteration (2S^2)."""
    return st["lu_factorizations"] * (2.0 * S ** 3 / 3.0) + st["newton_iters"] * 2.0 * S * S


def work_bytes(S, st):
    """Algorithmic HBM bytes of one instance: the dense Newton matrix written
    by the Jacobian assembly (8S^2) and written and read once by its LU
    (16S^2) per factorization, and the factors read once per Newton solve
    (8S^2)."""
    return st["lu_factorizations"] * 24.0 * S * S + st["newton_iters"] * 8.0 * S * S


def workload():
    from lassie_amd import synth
    w = WORKLOAD
    models = synth.uneven_batch(w["n"], w["seed_base"], sizes=w["sizes"], k_range=w["k_range"])
    t = synth.t_vector(0.0, w["t_end"], w["n_samples"])
    return models, t


def stratified(models, per_size):
    idx = []
    for S in WORKLOAD["sizes"]:
        idx += [i for i, m in enumerate(models) if m.S == S][:per_size]
    return idx


def cpu_legs(models, t, sample, prefix_loops):
    """Times the oracle and LSODA on the sample, one core, before any GPU work.
    Returns per-instance records; the oracle's large instances run a prefix."""
    from oracle import lsoda
    from oracle import oracle as O
    recs = []
    for i in sample:
        m = models[i]
        full = m.S <= 128
        kw = dict(CONTROLLER) if full else dict(CONTROLLER, max_iters=prefix_loops)
        t0 = time.perf_counter()
        _, st, rc, _, _ = O.solve(m, t, **kw)
        t_orc = time.perf_counter() - t0
        t0 = time.perf_counter()
        y_ls = lsoda.solve(m, t, rtol=1e-6)
        t_ls = time.perf_counter() - t0
        recs.append(dict(i=i, S=m.S, oracle_s=t_orc, oracle_full=full, oracle_bdf=int(st["bdf_steps"]),
                         oracle_samples=int(st["samples"]), lsoda_s=t_ls, lsoda_out=y_ls))
        say(f"  CPU S={m.S}: oracle {'full' if full else f'{prefix_loops}-iteration prefix'} {t_orc:.2f} s, "
            f"LSODA {t_ls:.2f} s")
    return recs


def project(models, per_size_time):
    """Workload time from per-size mean times: sum over sizes of count x mean."""
    tot = 0.0
    for S in WORKLOAD["sizes"]:
        n = sum(1 for m in models if m.S == S)
        tot += n * per_size_time[S]
    return tot


def differentiation_leg(B):
    """Gradients of the log loss w.r.t. log k for B parameterisations of the
    512 x 1024 network against targets from the hidden rates."""
    import dataclasses
    import numpy as np
    from lassie_amd import _lib, synth
    d = DIFF
    ms, hidden = synth.parameterisations(d["S"], d["R"], B, seed=d["seed"])
    t = synth.t_vector(0.0, d["t_end"], d["n_samples"])
    say(f"differentiation leg: {B} parameterisations of {d['S']}x{d['R']}, targets from the hidden rates")
    dms_h = [_lib.DeviceModel(dataclasses.replace(m, k=hidden[b])) for b, m in enumerate(ms)]
    tgt, _, _, _, _ = _lib.solve_batch(dms_h, t, raise_on_maxiter=False, **DIFF_CTRL)
    for dm in dms_h:
        dm.close()
    dms = [_lib.DeviceModel(m) for m in ms]
    t0 = time.perf_counter()
    outs, stats, fwd_ms, tape = _lib.forward_batch(dms, t, **DIFF_CTRL)
    eps = d["eps"]
    loss, W = 0.0, []
    for y, ys in zip(outs, tgt):
        r = np.log(y + eps) - np.log(ys + eps)
        loss += float(np.sum(r * r))
        W.append(2.0 * r / (y + eps))
    gks, _, rev_ms = tape.adjoint(W, allow_nonfinite=True)
    wall = time.perf_counter() - t0
    finite = [tape.status(i)[0] == 1 for i in range(B)]
    capped = sum(1 for i in range(B) if tape.status(i)[1] > 0)
    steps = [tape.info(i)[0] for i in range(B)]
    tape.close()
    for dm in dms:
        dm.close()
    say(f"  {B} gradients in {wall:.2f} s (forward {fwd_ms / 1e3:.2f} s, reverse {rev_ms / 1e3:.2f} s), "
        f"{sum(finite)} finite")
    return {"gradients_per_s": B / wall, "parameterisations": B, "wall_s": wall, "forward_ms": fwd_ms,
            "reverse_ms": rev_ms, "finite_fraction": sum(finite) / B,
            "parameterisations_with_capped_newton_steps": capped,
            "steps_per_parameterisation_mean": float(np.mean(steps)), "loss": loss,
            "config": dict(d, controller=DIFF_CTRL, hidden="k * U[0.5, 2], hidden seed 7"),
            "note": "a non-finite gradient comes from a BDF1 step whose modified Newton iteration ran all "
                    "10000 iterations without converging; the library reports it "
                    "(LASSIE_ERR_NONFINITE, lassie_adjoint_status); FD and oracle-adjoint agreement: GPU tests"}


def plain_1024():
    """One whole solve of the 1024 x 1024 stiff model (k logU[1e-6, 1e6], x0
    logU[1e-6, 1], seed 0, t in [0, 100], 200 samples) in the plain FP64 form
    (rtol 1e-6, BDF regrowth, at most 8 RKF rejections in a row): one launch of the single-model kernel, with its
    phase times and the dense-LU roofline (2/3 S^3 flops per factorization
    plus 2 S^2 per Newton iteration, against the FP64 peak)."""
    from lassie_amd import _lib, synth
    S = 1024
    m = synth.stiff(S, S, 0, (1e-6, 1e6))
    t = synth.t_vector(0.0, 100.0, 200)
    dm = _lib.DeviceModel(m)
    t0 = time.perf_counter()
    _, _, st, rc, _ = dm.solve(t, raise_on_maxiter=False, rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
    wall = time.perf_counter() - t0
    dm.close()
    lu = max(st["lu_factorizations"], 1)
    flops = 2.0 / 3.0 * S ** 3 * st["lu_factorizations"] + 2.0 * S * S * st["newton_iters"]
    achieved = flops / (st["kernel_ms"] / 1e3) / 1e12
    res = {"wall_s": wall, "kernel_ms": st["kernel_ms"], "rc": rc, "loop_iters": st["loop_iters"],
           "bdf_steps": st["bdf_steps"], "newton_iters": st["newton_iters"],
           "lu_factorizations": st["lu_factorizations"],
           "phase_ms": {"control_rkf": st["phase_ms"][0], "jacobian": st["phase_ms"][1], "lu": st["phase_ms"][2],
                        "newton_solves": st["phase_ms"][3]},
           "lu_ms_per_factorization": st["phase_ms"][2] / lu,
           "newton_us_per_iteration": 1e3 * st["phase_ms"][3] / max(st["newton_iters"], 1),
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}}
    traffic = os.path.join(ROOT, "profiles", "traffic_solve_1024.json")
    if os.path.exists(traffic):
        with open(traffic) as f:
            res["roofline"]["traffic"] = json.load(f).get("bytes_per_launch")
    say(f"  1024x1024 plain solve: {st['kernel_ms'] / 1e3:.2f} s, LU {res['lu_ms_per_factorization']:.3f} ms, "
        f"Newton iteration {res['newton_us_per_iteration']:.1f} us")
    return res


PRECISION_FORMS = ("fp64", "fp32_refine", "bf16_refine", "fp32_unrefined", "bf16_unrefined")


def precision_form_4096(cap, forms=PRECISION_FORMS):
    """The FP32-refinement configuration: the 4096 x 4096 stiff model
    (k logU[1e-6, 1e6], x0 logU[1e-6, 1], seed 0, t in [0, 100], 200 samples)
    in the plain FP64 form and in the single-model precision forms (FP32 grid
    LU with 8-column panels, FP64 refinement of every Newton correction to a
    componentwise backward error with the species' atol as floor; the same
    with the trailing updates on the BF16 matrix cores, 30 sweeps; and both
    unrefined, refine_max_iter = 0: one FP32 solve per Newton iteration, the
    modified-Newton iteration's FP64 residual being the only refinement), over
    the first `cap` loop iterations of the same controller; phase times per
    factorization and per Newton iteration, and the error of the final state
    against the FP64 form."""
    import numpy as np
    from lassie_amd import _lib, synth
    m = synth.stiff(4096, 4096, 0, (1e-6, 1e6))
    t = synth.t_vector(0.0, 100.0, 200)
    ctrl = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
    dm = _lib.DeviceModel(m)
    res, finals = {}, {}
    table = {"fp64": {}, "fp32_refine": dict(precision=_lib.PREC_FP32_REFINE),
             "bf16_refine": dict(precision=_lib.PREC_BF16_REFINE, refine_max_iter=30),
             "fp32_unrefined": dict(precision=_lib.PREC_FP32_REFINE, refine_max_iter=0),
             "bf16_unrefined": dict(precision=_lib.PREC_BF16_REFINE, refine_max_iter=0)}
    forms = ["fp64"] + [f for f in forms if f != "fp64"]
    for name in forms:
        kw = table[name]
        say(f"  4096x4096 {name}, first {cap} loop iterations")
        t0 = time.perf_counter()
        _, yf, st, rc, _ = dm.solve(t, raise_on_maxiter=False, max_iters=cap, **ctrl, **kw)
        wall = time.perf_counter() - t0
        finals[name] = yf
        lu = max(st["lu_factorizations"], 1)
        res[name] = {"wall_s": wall, "kernel_ms": st["kernel_ms"], "bdf_steps": st["bdf_steps"],
                     "newton_iters": st["newton_iters"], "lu_factorizations": st["lu_factorizations"],
                     "lu_ms_per_factorization": st["phase_ms"][2] / lu,
                     "newton_ms_per_iteration": st["phase_ms"][3] / max(st["newton_iters"], 1),
                     "refine_sweeps": st["refine_iters"], "refine_fallbacks": st["refine_fallbacks"],
                     "groups": st["groups"]}
        say(f"  {name}: {wall:.1f} s, LU {res[name]['lu_ms_per_factorization']:.2f} ms, Newton iteration "
            f"{res[name]['newton_ms_per_iteration']:.3f} ms")
    dm.close()
    a = finals["fp64"]
    for name in forms[1:]:
        res[name]["max_rel_err_final_state_vs_fp64"] = float(np.max(np.abs(finals[name] - a) / (np.abs(a) + m.atol)))
    if "fp32_refine" in res:
        res["max_rel_err_final_state_vs_fp64"] = res["fp32_refine"]["max_rel_err_final_state_vs_fp64"]
    res["loop_iterations"] = cap
    return res


def calibrate_copy():
    """Three copies of a 256 MiB FP64 buffer (each reads and writes 262144 KiB):
    the known traffic that scales the HBM counters to bytes."""
    import torch
    n = 256 * 1024 * 1024 // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    say(f"copied 3 x 256 MiB: {bool(b[123] == a[123])}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes first (at most one is run)")
    ap.add_argument("--budget", type=float, default=420.0,
                    help="wall-clock seconds for the whole run; timed steps stop when the next one would "
                         "not fit, and the line says how many ran")
    ap.add_argument("--plain-per-size", type=int, default=1, help="plain-path instances per size")
    ap.add_argument("--cpu-per-size", type=int, default=1, help="CPU-baseline instances per size")
    ap.add_argument("--cpu-prefix", type=int, default=1500,
                    help="loop iterations of the oracle on S >= 256 sample instances")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--diff-batch", type=int, default=DIFF["B"],
                    help="parameterisations of the differentiation leg (0 skips it)")
    ap.add_argument("--precision-4096", type=int, default=0, metavar="N",
                    help="also compare the plain and precision forms on the 4096x4096 model over its first N "
                         "loop iterations (not part of the default run)")
    ap.add_argument("--precision-forms", default=",".join(PRECISION_FORMS),
                    help="comma-separated forms of --precision-4096 (fp64 always runs, as the reference)")
    ap.add_argument("--precision-batch", action="store_true",
                    help="also time the batched precision form (FP32 LU + FP64 refinement) on the "
                         "workload and report its error against the chosen form")
    ap.add_argument("--plain-1024", action="store_true",
                    help="also run one whole plain FP64 solve of the 1024x1024 stiff model (about 30 s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy one 256 MiB FP64 buffer into another three times (a kernel of known "
                         "traffic, to scale the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: instances are independent, "
              "so several GPUs would each run their own share of the batch in separate processes, with no "
              "collective in the solve path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    say("generating the 2048-instance workload")
    models, t = workload()
    cpu_sample = stratified(models, args.cpu_per_size)
    cpu = None
    if not args.no_cpu:
        say(f"CPU baselines on {len(cpu_sample)} instances (one per size unless asked), one core")
        cpu = cpu_legs(models, t, cpu_sample, args.cpu_prefix)

    from lassie_amd import _lib
    say("uploading models")
    dms = [_lib.DeviceModel(m) for m in models]
    plan = _lib.BatchPlan(dms, t, **CONTROLLER)
    variant, tile_rows, trial_ms = plan.info()
    say(f"plan prepared (variant {variant}, trials {[round(x, 1) for x in trial_ms]} ms)")

    finished = threading.Event()

    def follow(t_run):
        last, quiet = 0, 0
        while not finished.wait(5.0):
            done = plan.progress()
            quiet += 1
            if done != last or quiet >= 6:
                last, quiet = done, 0
                say(f"  {done} of {len(models)} instances finished ({time.perf_counter() - t_run:.1f} s)")

    def run_once():
        finished.clear()
        t_run = time.perf_counter()
        th = threading.Thread(target=follow, args=(t_run,), daemon=True)
        th.start()
        try:
            res = plan.run(raise_on_maxiter=False)
        finally:
            finished.set()
            th.join()
        return res, time.perf_counter() - t_run

    # One warmup pass at most: a pass is tens of seconds of deterministic
    # work, so further warmups change nothing and the budget goes to timed
    # passes instead.
    plain_reserve = 45.0 * args.plain_per_size + 15.0 + (45.0 if args.diff_batch > 0 else 0.0)
    warm, first_pass = 0, None
    if args.warmup > 0:
        say("warmup pass (instances handed out largest S first)")
        _, first_pass = run_once()
        warm = 1
        say(f"  warmup pass: {first_pass:.2f} s; later passes hand instances out longest measured time first")
    say("timed passes")
    walls, kernel_ms, res = [], [], None
    while len(walls) < args.steps:
        if walls and elapsed() + max(walls) + plain_reserve > args.budget:
            say(f"budget: stopping after {len(walls)} timed pass(es)")
            break
        res, w = run_once()
        walls.append(w)
        kernel_ms.append(plan.timing()[2])
        say(f"  pass {len(walls)}: {w:.2f} s")
    outs, ys, stats, dev_ms, rc = res
    n_wave, n_wg, _ = plan.timing()
    steps = len(walls)
    wall = float(np.sum(walls))
    value = len(models) * steps / wall
    capped = [i for i, s in enumerate(stats) if s["samples"] < t.size]

    # roofline of the workgroup-level kernel, from its own events
    wg_idx = [i for i, m in enumerate(models) if m.S > 64]
    wv_idx = [i for i, m in enumerate(models) if m.S <= 64]
    wg_flops = sum(work_flops(models[i].S, stats[i]) for i in wg_idx)
    wv_flops = sum(work_flops(models[i].S, stats[i]) for i in wv_idx)
    wg_ms = float(np.mean([k[1] for k in kernel_ms]))
    wv_ms = float(np.mean([k[0] for k in kernel_ms]))
    achieved = wg_flops / (wg_ms / 1e3) / 1e12
    wg_bytes = sum(work_bytes(models[i].S, stats[i]) for i in wg_idx)

    # plain path on a stratified sample of the same instances, bit-exact check
    plain_idx = stratified(models, args.plain_per_size)
    say(f"plain path on {len(plain_idx)} instances")
    plain_t, mismatches = {}, 0
    for i in plain_idx:
        t0 = time.perf_counter()
        out_p, y_p, st_p, rc_p, _ = dms[i].solve(t, raise_on_maxiter=False, **CONTROLLER)
        dt = time.perf_counter() - t0
        plain_t.setdefault(models[i].S, []).append(dt)
        n = int(st_p["samples"])
        same = (np.array_equal(out_p[:n], outs[i][:n]) and np.array_equal(y_p, ys[i])
                and st_p["newton_iters"] == stats[i]["newton_iters"])
        mismatches += 0 if same else 1
        say(f"  plain S={models[i].S}: {dt:.2f} s, {'bit-identical' if same else 'DIFFERS'}")
    plain_mean = {S: float(np.mean(v)) for S, v in plain_t.items()}
    plain_rate = len(models) / project(models, plain_mean) if len(plain_mean) == len(WORKLOAD["sizes"]) else None

    per_size = {}
    for S in WORKLOAD["sizes"]:
        idx = [i for i, m in enumerate(models) if m.S == S]
        ph = np.array([stats[i]["phase_ms"] for i in idx])
        per_size[str(S)] = {
            "instances": len(idx), "capped": sum(1 for i in idx if i in capped),
            "instance_ms_mean": float(ph.sum(1).mean()), "instance_ms_max": float(ph.sum(1).max()),
            "lu_us": float(1e3 * ph[:, 2].sum() / max(1, sum(stats[i]["lu_factorizations"] for i in idx))),
            "newton_iter_us": float(1e3 * ph[:, 3].sum() / max(1, sum(stats[i]["newton_iters"] for i in idx))),
            "plain_s_mean": plain_mean.get(S)}

    line = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2048 uneven stiff mass-action instances (S in {16..512}, R = 2S, seeds "
                               "1000..3047, k logU[1e-4, 1e4]), t = [0, 50], 100 samples, rtol 1e-6; "
                               "batched FP64 plan vs plain FP64 single-model solves",
                   "instances": len(models), "sizes": list(WORKLOAD["sizes"]), "controller": CONTROLLER,
                   "form": "batched FP64 plan: workgroup-level kernel (S > 64) and wave-level kernel "
                           "(S <= 64) concurrently, hipGraph replay"},
        "chosen_vs_plain": {
            "chosen_solves_per_s": value, "plain_solves_per_s_projected": plain_rate,
            "speedup": value / plain_rate if plain_rate else None,
            "max_rel_err_vs_plain": (0.0 if mismatches == 0 else None) if plain_idx else None,
            "plain_sample": {"instances": len(plain_idx), "per_size": args.plain_per_size,
                             "bit_identical": mismatches == 0,
                             "note": "plain rate = 2048 / sum over sizes of (instances x the sample's mean "
                                     "plain solve time)"}},
        "capped": {"instances": len(capped), "max_newton": CONTROLLER["max_newton"],
                   "note": "instances stopped by the per-instance Newton-iteration cap (same cap and same "
                           "stopping point in every form and in the oracle)"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "lassie::batch_kernel (workgroup-level, S > 64 instances)",
                     "flops_per_launch": wg_flops, "kernel_ms": wg_ms, "algorithmic_bytes_per_launch": wg_bytes,
                     "wave_kernel": {"instances": n_wave, "ms": wv_ms, "flops": wv_flops,
                                     "tflops": wv_flops / (wv_ms / 1e3) / 1e12 if wv_ms > 0 else None}},
        "per_size": per_size, "device_ms_per_step": dev_ms, "walls_s": walls,
        "schedule": {"first_pass_s": first_pass,
                     "wg_kernel_workgroups": int(stats[0]["groups"]) if stats else None,
                     "note": "the plan's first run hands instances out largest S first on every CU (first_pass_s, "
                             "the warmup pass); every run records each instance's device time, and the next run "
                             "hands them out longest first and leaves the wave-level kernel as many CUs as balance "
                             "the two kernels' measured work, so the timed passes run the measured-cost schedule"},
    }
    traffic = os.path.join(ROOT, "profiles", "traffic_batch_c4.json")
    if os.path.exists(traffic):
        with open(traffic) as f:
            tr = json.load(f)
        line["roofline"]["traffic"] = tr.get("bytes_per_launch")
        line["roofline"]["traffic_GBps"] = tr.get("bytes_per_launch") / (wg_ms / 1e3) / 1e9
        line["roofline"]["traffic_over_algorithmic"] = tr.get("bytes_per_launch") / wg_bytes

    if cpu is not None:
        # oracle: full solves below 256 species; larger ones scaled from their
        # prefix by the instance's full BDF-step count in the GPU run
        orc = {}
        for r in cpu:
            full_s = r["oracle_s"] if r["oracle_full"] else \
                r["oracle_s"] * stats[r["i"]]["bdf_steps"] / max(1, r["oracle_bdf"])
            orc.setdefault(r["S"], []).append(full_s)
        orc_mean = {S: float(np.mean(v)) for S, v in orc.items()}
        orc_rate = len(models) / project(models, orc_mean)
        lso_mean = {}
        for r in cpu:
            lso_mean.setdefault(r["S"], []).append(r["lsoda_s"])
        lso_mean = {S: float(np.mean(v)) for S, v in lso_mean.items()}
        lso_rate = len(models) / project(models, lso_mean)
        errs = []
        for r in cpu:
            i = r["i"]
            if stats[i]["samples"] == t.size:
                y = r["lsoda_out"]
                errs.append(np.abs(outs[i] - y) / (np.abs(y) + models[i].atol[None, :]))
        allerr = np.concatenate([e.ravel() for e in errs]) if errs else np.zeros(1)
        line["cpu_baseline"] = {
            "value": orc_rate, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"the oracle (FP64 CPU restatement of the same RKF45/BDF1 loop, same controller and cap) "
                      f"on the first {args.cpu_per_size} instance(s) of each size: full solves for S <= 128, "
                      f"the first {args.cpu_prefix} loop iterations for S = 256 and 512 scaled by the "
                      f"instance's full BDF-step count; workload rate = 2048 / sum of (instances x mean)",
            "per_size_s": orc_mean}
        line["cpu_lsoda"] = {
            "value": lso_rate, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "LSODA (scipy odeint, the reference's CPU method; rtol 1e-6, atol 1e-12) on the same "
                      "instances, projected the same way; a different algorithm (variable-order BDF)",
            "per_size_s": lso_mean,
            "rel_err_vs_chosen_median": float(np.median(allerr)), "rel_err_vs_chosen_max": float(allerr.max())}

    if args.precision_batch and elapsed() + 2 * (wall / steps) < args.budget:
        say("batched precision form (FP32 LU + FP64 refinement)")
        p32 = _lib.BatchPlan(dms, t, precision=_lib.PREC_FP32_REFINE, **CONTROLLER)
        t0 = time.perf_counter()
        o32, _, s32, _, _ = p32.run(raise_on_maxiter=False)
        w32 = time.perf_counter() - t0
        p32.close()
        rel = [float((np.abs(a - b) / (np.abs(b) + m.atol[None, :])).max())
               for a, b, m in zip(o32, outs, models)]
        line["precision_batch"] = {"solves_per_s": len(models) / w32, "max_rel_err_vs_fp64": float(np.max(rel)),
                                   "median_instance_max_rel_err": float(np.median(rel)),
                                   "fallbacks": int(sum(s["refine_fallbacks"] for s in s32))}
    plan.close()
    for dm in dms:
        dm.close()
    if args.diff_batch > 0:
        line["differentiation_form"] = differentiation_leg(args.diff_batch)
    if args.plain_1024:
        say("plain FP64 solve of the 1024x1024 model")
        line["plain_1024"] = plain_1024()
    if args.precision_4096 > 0:
        say(f"precision form on 4096x4096, {args.precision_4096} loop iterations")
        line["precision_form_4096"] = precision_form_4096(args.precision_4096, args.precision_forms.split(","))
    line["bench_wall_s"] = elapsed()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
