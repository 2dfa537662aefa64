"""Batched persistent form vs the oracle.

FP64: every instance of the batch must equal the oracle bit for bit (same
algorithm and operation order as the single-model path).  FP32 / BF16 LU with
FP64 refinement: on well-conditioned Newton matrices, the FP64 form's
decisions and a 1e-4 bound on |a - b| / (|b| + atol); on stiff instances the
relative error and the refinement work are reported (the loop's accept/reject
decisions are discontinuous, so no bound holds for every model).
The batched plan (graph replay, autotuned variant) equals the batched form.
"""
import dataclasses

import numpy as np
import pytest

from lassie_amd import _lib, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)


def _batch(seed_base, n, sizes):
    rng = np.random.default_rng(seed_base)
    pick = rng.integers(0, len(sizes), size=n)
    return [synth.stiff(int(sizes[p]), 2 * int(sizes[p]), seed_base + i, (1e-4, 1e4))
            for i, p in enumerate(pick)]


def test_batch_fp64_bit_exact(gpu_lib):
    models = _batch(2000, 12, (16, 32, 48, 64))
    t = synth.t_vector(0, 50, 100)
    print(f"[batch] fp64 {len(models)} instances sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    outs, ys, stats, ms, rc = gpu_lib.solve_batch(dms, t, max_iters=2_000_000, **BENCH)
    assert rc == 0
    for m, out, y, st in zip(models, outs, ys, stats):
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, max_iters=2_000_000, **BENCH)
        assert o_rc == 0
        for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
                    "lu_failures", "samples"):
            assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
        np.testing.assert_array_equal(out, o_out)
        np.testing.assert_array_equal(y, o_y)
    for dm in dms:
        dm.close()


def test_batch_single_model_agrees_with_single_path(gpu_lib):
    m = synth.stiff(96, 96, 5)
    t = synth.t_vector(0, 100, 200)
    dm = gpu_lib.DeviceModel(m)
    out1, y1, st1, rc1, _ = dm.solve(t, raise_on_maxiter=False, max_iters=300, **BENCH)
    outs, ys, stats, ms, rc = gpu_lib.solve_batch([dm], t, raise_on_maxiter=False, max_iters=300, **BENCH)
    n = int(st1["samples"])
    np.testing.assert_array_equal(outs[0][:n], out1[:n])
    np.testing.assert_array_equal(ys[0], y1)
    dm.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16"])
def test_batch_mixed_precision_mild_steps(gpu_lib, prec):
    """FP32 / BF16 factors with FP64 refinement where the Newton matrices are
    well conditioned (k logU[1e-4, 1e2], BDF step 1e-3, t in [0, 1]):
    refinement meets its componentwise target (floor: the species' atol)
    in every solve, the controller takes the FP64 form's decisions, and
    every instance stays within 1e-4 of it, |a - b| / (|b| + atol) (6.2e-5 at
    worst measured: the step-size feedback of the RKF45 controller amplifies
    1e-16 differences of the corrections)."""
    models = [synth.stiff(S, 2 * S, 7000 + i, (1e-4, 1e2)) for i, S in enumerate((16, 48, 100, 200, 333, 512))]
    t = synth.t_vector(0, 1, 11)
    kw = dict(be_step=1e-3, max_iters=2_000_000, **BENCH)
    print(f"[batch] precision {prec} mild sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    ref, _, ref_st, _, rc0 = gpu_lib.solve_batch(dms, t, **kw)
    outs, _, stats, _, rc = gpu_lib.solve_batch(dms, t, precision=prec, refine_max_iter=10, **kw)
    assert rc0 == 0 and rc == 0
    for m, a, b, st, rs in zip(models, outs, ref, stats, ref_st):
        rel = float((np.abs(a - b) / (np.abs(b) + m.atol[None, :])).max())
        print(f"[batch] S={m.S} rel={rel:.2e} refine={st['refine_iters']} "
              f"fallbacks={st['refine_fallbacks']} bdf={st['bdf_steps']}", flush=True)
        assert st["refine_iters"] >= st["newton_iters"] > 0 and st["refine_fallbacks"] == 0
        for k in ("loop_iters", "bdf_steps", "newton_iters", "nan_repairs"):
            assert st[k] == rs[k], k
        assert rel <= 1e-4
    for dm in dms:
        dm.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16"])
def test_batch_mixed_precision_stiff(gpu_lib, prec):
    """The experimental precision forms on uneven stiff instances of the
    2048-instance workload's generator (S in {16, 32, 64, 128}, k logU[1e-4,
    1e4], t in [0, 50]) against the FP64 oracle: every instance completes
    with finite samples, refinement converges (at most 1% of BDF1 steps fall
    back to the FP64 LU), and the median relative error of each instance's
    samples is at most 5e-2.  The maximum is printed, not asserted: the
    loop's accept/reject and Newton-count decisions are discontinuous, so
    corrections that differ at 1e-14 lead to different step sequences (the
    oracle's own precision form shows per-instance medians up to 1.6e-2 and
    maxima up to 1e7 on species near zero on these instances)."""
    models = _batch(3000, 10, (16, 32, 64, 128))
    t = synth.t_vector(0, 50, 100)
    kw = dict(max_newton=50000, **BENCH)
    print(f"[batch] precision {prec} sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    outs, _, stats, _, rc = gpu_lib.solve_batch(dms, t, precision=prec, raise_on_maxiter=False, **kw)
    checked = 0
    for m, a, st in zip(models, outs, stats):
        ref, o_st, o_rc, _, _ = O.solve(m, t, **kw)
        if o_rc or st["samples"] < t.size:
            continue  # capped: not a complete trajectory in one of the forms
        checked += 1
        assert np.all(np.isfinite(a))
        assert st["refine_iters"] > 0 and st["refine_fallbacks"] <= 0.01 * st["bdf_steps"] + 1
        rel = np.abs(a - ref) / (np.abs(ref) + m.atol[None, :])
        print(f"[batch] S={m.S} median rel={np.median(rel):.2e} max rel={rel.max():.2e} "
              f"within 1e-6: {(rel <= 1e-6).mean():.2f} refine={st['refine_iters']} "
              f"fallbacks={st['refine_fallbacks']}", flush=True)
        assert np.median(rel) <= 5e-2
    assert checked >= len(models) // 2
    for dm in dms:
        dm.close()


def test_batch_plan_replays_the_batched_form(gpu_lib):
    """lassie_batch_plan_*: the graph replay gives the batched form's results
    bit for bit, twice in a row, and picks up new rates; the autotuner reports
    its choice; the finished-instance count covers every instance after each run."""
    models = synth.uneven_batch(10, 4242, sizes=(16, 48, 100, 200))
    t = synth.t_vector(0, 2, 9)
    kw = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
    print("[plan] 10 uneven instances", flush=True)
    dms = [_lib.DeviceModel(m) for m in models]
    ref, ref_y, ref_st, _, _ = _lib.solve_batch(dms, t, **kw)
    plan = _lib.BatchPlan(dms, t, **kw)
    variant, rt, trial = plan.info()
    assert variant in range(_lib.BATCH_VARIANTS) and rt in (1, 2, 4)
    assert plan.progress() == 0
    for _ in range(2):
        outs, ys, st, ms, rc = plan.run()
        assert rc == 0 and ms > 0
        assert plan.progress() == len(models)
        for a, b in zip(outs, ref):
            assert np.array_equal(a, b)
        assert [s["bdf_steps"] for s in st] == [s["bdf_steps"] for s in ref_st]
    # new rates: the next replay integrates them
    k2 = models[3].k * 1.5
    dms[3].set_rates(k2)
    outs, _, _, _, _ = plan.run()
    m3 = dataclasses.replace(models[3], k=k2)
    o_out, _, o_rc, _, _ = O.solve(m3, t, **kw)
    assert np.array_equal(outs[3], o_out)
    plan.close()


_TUNE_CHILD = r'''
import json, sys
sys.path.insert(0, sys.argv[1])
from lassie_amd import _lib, synth
models = synth.uneven_batch(6, 4444, sizes=(100, 200))
t = synth.t_vector(0, 1, 5)
dms = [_lib.DeviceModel(m) for m in models]
plan = _lib.BatchPlan(dms, t, rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
print(json.dumps(plan.info()))
plan.close()
'''


def test_plan_autotuner_choice_persists(gpu_lib, tmp_path):
    """The autotuner stores its choice in the cache file (valid JSON, keyed by
    device and workload shape); a new process with the same file reuses it
    and runs no trials (trial times all zero)."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.json"
    env = dict(os.environ, LASSIE_TUNE_CACHE=str(cache))
    runs = []
    for _ in range(2):
        print("[plan] autotuner child process", flush=True)
        r = subprocess.run([sys.executable, "-c", _TUNE_CHILD, root], env=env, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    (v1, _, ms1), (v2, _, ms2) = runs
    assert max(ms1) > 0 and ms2 == [0.0] * _lib.BATCH_VARIANTS and v1 == v2
    doc = json.loads(cache.read_text())
    assert doc["format"] == "lassie-tune-1"
    (key, entry), = doc["entries"].items()
    assert key.startswith("batch_variant/") and "/smax200/n6/prec0" in key
    assert entry["choice"] == v1 and len(entry["trial_ms"]) == _lib.BATCH_VARIANTS


def test_batch_fp64_bit_exact_large_c4_sizes(gpu_lib):
    """Instances of the 2048-instance workload's generator at S = 128, 256
    and 512 (R = 2S, k logU[1e-4, 1e4], t in [0, 50], 100 samples) on the
    workgroup-level kernel against the oracle, bit for bit: S = 128 over the
    whole horizon (with the bench's Newton cap), S = 256 and 512 over their
    first 1500 loop iterations."""
    models = synth.uneven_batch(6, 1000, sizes=(128, 256, 512))
    t = synth.t_vector(0.0, 50.0, 100)
    for cap in ((0, (128,)), (1500, (256, 512))):
        ms = [m for m in models if m.S in cap[1]]
        if not ms:
            continue
        kw = dict(BENCH, max_newton=50000, max_iters=cap[0])
        print(f"[batch] large C4 sizes {[m.S for m in ms]} max_iters={cap[0]}", flush=True)
        dms = [gpu_lib.DeviceModel(m) for m in ms]
        outs, ys, stats, _, rc = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
        for dm in dms:
            dm.close()
        for m, out, y, st in zip(ms, outs, ys, stats):
            o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
            for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
                        "lu_failures", "samples"):
                assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
            n = int(o_st["samples"])
            np.testing.assert_array_equal(out[:n], o_out[:n])
            np.testing.assert_array_equal(y, o_y)


def test_batch_fp64_bit_exact_odd_sizes(gpu_lib):
    """Odd and even S above the wave-level limit in one batch whose largest
    instance is odd (S in {66, 97, 130, 131}): the per-slot workspace stride
    is then odd, so factors sit at 8-byte as well as 16-byte aligned
    addresses, and the Newton solve runs with both of its load widths (8-byte
    loads for odd S or unaligned factors, 16-byte loads otherwise).  Every
    instance against the oracle, bit for bit, over its first 3000 loop
    iterations (k logU[1e-4, 1e4], t in [0, 20], 40 samples)."""
    models = synth.uneven_batch(8, 7100, sizes=(66, 97, 130, 131))
    models[0] = synth.stiff(131, 262, 7100, (1e-4, 1e4))
    t = synth.t_vector(0.0, 20.0, 40)
    kw = dict(BENCH, max_newton=50000, max_iters=3000)
    print(f"[batch] odd sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    outs, ys, stats, _, rc = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
    for dm in dms:
        dm.close()
    newton = 0
    for m, out, y, st in zip(models, outs, ys, stats):
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
        for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
                    "lu_failures", "samples"):
            assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
        n = int(o_st["samples"])
        np.testing.assert_array_equal(out[:n], o_out[:n])
        np.testing.assert_array_equal(y, o_y)
        newton += st["newton_iters"]
    assert newton > 0  # the Newton solve ran


def _singular_at(S, col, h=0.01):
    """S species: species `col` is autocatalytic (A -> 2A, k = 1/h, so the
    Newton matrix entry 1 - hk is 0 at BDF step h and nothing else touches
    column `col`); the others decay and convert into each other."""
    from tests import cases
    reactions, k = [], []
    for s in range(S):
        if s == col:
            reactions.append(({s: 1}, {s: 2}))
            k.append(1.0 / h)
        else:
            reactions.append(({s: 1}, {}))
            k.append(0.5 + 0.01 * s)
    rng = np.random.default_rng(3)
    others = [s for s in range(S) if s != col]
    for _ in range(2 * S):
        a, b = rng.choice(others, 2, replace=False)
        reactions.append(({int(a): 1}, {int(b): 1}))
        k.append(float(10 ** rng.uniform(-2, 2)))
    return cases.from_reactions(S, reactions, k, np.linspace(0.1, 1.0, S))


def test_batch_fp64_bit_exact_blocked_lu_sizes(gpu_lib):
    """The workgroup-level LU factors matrices with S % 16 == 0 and S > 256 in
    64-column blocks.  Sizes that are not multiples of 64 (272 ... 496) end
    with a 16-, 32- or 48-column block, and their trailing regions split
    into unequal LDS chunks of row tiles; each instance of the uneven
    generator against the oracle, bit for bit, over its first 2000 loop
    iterations.  A 272-species model whose first BDF step meets a zero pivot
    at column 200 (inside the fourth block; bench controller over the whole
    horizon, faithful controller over 200 loop iterations) checks the
    LU-failure path of the blocked factorization."""
    sizes = (272, 304, 352, 400, 432, 496)
    models = [synth.uneven_batch(1, 7300 + S, sizes=(S,))[0] for S in sizes]
    t = synth.t_vector(0.0, 50.0, 100)
    kw = dict(BENCH, max_newton=50000, max_iters=2000)
    print(f"[batch] blocked-LU sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    outs, ys, stats, _, rc = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
    for dm in dms:
        dm.close()
    for m, out, y, st in zip(models, outs, ys, stats):
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
        for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
                    "lu_failures", "samples"):
            assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
        n = int(o_st["samples"])
        np.testing.assert_array_equal(out[:n], o_out[:n])
        np.testing.assert_array_equal(y, o_y)
        assert st["bdf_steps"] > 0
    m = _singular_at(272, 200)
    t = synth.t_vector(0.0, 1.0, 11)
    for opts, cap in ((BENCH, 20000), (dict(), 200)):
        print(f"[batch] zero pivot at column 200 of 272, opts={opts} max_iters={cap}", flush=True)
        dm = gpu_lib.DeviceModel(m)
        outs, ys, stats, _, rc = gpu_lib.solve_batch([dm], t, raise_on_maxiter=False, max_iters=cap, **opts)
        dm.close()
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, max_iters=cap, **opts)
        assert o_st["lu_failures"] >= 1
        for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
                    "lu_failures", "samples"):
            assert stats[0][key] == o_st[key], (key, stats[0][key], o_st[key])
        n = int(o_st["samples"])
        np.testing.assert_array_equal(outs[0][:n], o_out[:n])
        np.testing.assert_array_equal(ys[0], o_y)


def test_batch_hub_rows_bit_exact(gpu_lib):
    """Jacobian rows with more distinct columns than the register-accumulated
    form holds (the hub species of cases.hub) take the read-modify-write
    form: both kernels against the oracle, bit for bit."""
    from tests import cases
    models = [cases.hub(20), cases.hub(40)]  # S = 41 (wave-level), 81 (workgroup-level)
    t = synth.t_vector(0.0, 5.0, 21)
    kw = dict(BENCH, max_iters=200_000)
    print(f"[batch] hub models {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    outs, ys, stats, _, rc = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
    for dm in dms:
        dm.close()
    for m, out, y, st in zip(models, outs, ys, stats):
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
        assert o_st["bdf_steps"] > 0
        for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs", "lu_failures",
                    "samples"):
            assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
        n = int(o_st["samples"])
        np.testing.assert_array_equal(out[:n], o_out[:n])
        np.testing.assert_array_equal(y, o_y)


def test_batch_overflowing_rate_prefix(gpu_lib):
    """The workgroup-level kernel on a Newton matrix with infinite entries
    (cases.overflowing_rate, S = 300): the same panels, repairs and
    trajectory as the oracle over the first 30 loop iterations."""
    from tests import cases
    m = cases.overflowing_rate()
    t = synth.t_vector(0.0, 100.0, 200)
    kw = dict(BENCH, max_iters=30, newton_iter=100)
    print(f"[batch] overflowing rate S={m.S}", flush=True)
    dm = gpu_lib.DeviceModel(m)
    outs, ys, stats, _, rc = gpu_lib.solve_batch([dm], t, raise_on_maxiter=False, **kw)
    dm.close()
    o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
    assert o_st["nan_repairs"] >= 10
    for key in ("loop_iters", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs", "lu_failures", "samples"):
        assert stats[0][key] == o_st[key], (key, stats[0][key], o_st[key])
    n = int(o_st["samples"])
    np.testing.assert_array_equal(outs[0][:n], o_out[:n])
    np.testing.assert_array_equal(ys[0], o_y)



@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16"])
def test_batch_unrefined_is_refinement_with_an_unbounded_target(gpu_lib, prec):
    """refine_max_iter = 0 in the batched precision forms (one FP32 solve per
    Newton iteration, x = 0 + (double)d) equals one refinement sweep whose
    target every finite correction meets (refine_tol = 1e300), bit for bit,
    counters included, at sizes on the 16-column LU (48, 200) and the
    blocked one (512); first 60 loop iterations, BDF steps of 1e-3 on
    well-conditioned matrices (k logU[1e-4, 1e2]).  (A correction that
    overflows gives a NaN residual, which the refined form does not accept.)"""
    models = [synth.stiff(S, 2 * S, 7100 + i, (1e-4, 1e2)) for i, S in enumerate((48, 200, 512))]
    t = synth.t_vector(0, 1, 11)
    kw = dict(max_iters=60, be_step=1e-3, precision=prec, **BENCH)
    print(f"[batch] unrefined precision {prec} sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    a, ya, sa, _, _ = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, refine_max_iter=0, **kw)
    b, yb, sb, _, _ = gpu_lib.solve_batch(dms, t, raise_on_maxiter=False, refine_tol=1e300, refine_max_iter=4, **kw)
    for i, m in enumerate(models):
        for k in ("loop_iters", "bdf_steps", "newton_iters", "lu_failures", "nan_repairs", "samples",
                  "refine_iters", "refine_fallbacks"):
            assert sa[i][k] == sb[i][k], (m.S, k)
        # one FP32 solve per Newton iteration, except in steps moved to the
        # FP64 LU by a zero FP32 pivot
        assert sa[i]["refine_iters"] > 0
        assert sa[i]["refine_fallbacks"] > 0 or sa[i]["refine_iters"] == sa[i]["newton_iters"]
        np.testing.assert_array_equal(ya[i], yb[i])
        np.testing.assert_array_equal(a[i], b[i])
    for dm in dms:
        dm.close()
