"""Batched persistent form vs the oracle.

FP64: every instance of the batch must equal the oracle bit for bit (same
algorithm and operation order as the single-model path).  FP32 / BF16 LU with
FP64 refinement: the relative error against the FP64 form,
|a - b| / (|b| + atol), and the refinement work are reported (the loop's
accept/reject decisions are discontinuous, so no bound holds for every model).
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
def test_batch_mixed_precision(gpu_lib, prec):
    """FP32 / BF16 factors with FP64 refinement on uneven stiff instances.

    Where the refinement meets its componentwise backward-error target the
    Newton corrections equal the FP64 ones to rounding.  Checked: every
    instance completes with finite samples, refinement work is reported, and
    at least one instance follows the FP64 form's decision sequence exactly.
    The relative error |a - b| / (|b| + atol) against the FP64 form is printed
    per instance (not asserted: on these stiff models a near-zero species can
    change sign between two accurate solves and flip a repair decision).
    """
    models = _batch(3000, 10, (16, 32, 64, 128))
    t = synth.t_vector(0, 50, 100)
    print(f"[batch] precision {prec} sizes {[m.S for m in models]}", flush=True)
    dms = [gpu_lib.DeviceModel(m) for m in models]
    ref, _, ref_st, _, rc0 = gpu_lib.solve_batch(dms, t, max_iters=2_000_000, **BENCH)
    outs, _, stats, _, rc = gpu_lib.solve_batch(dms, t, max_iters=2_000_000, precision=prec, **BENCH)
    assert rc0 == 0 and rc == 0
    same = 0
    worst_same = 0.0
    for m, a, b, st, rs in zip(models, outs, ref, stats, ref_st):
        assert st["samples"] == t.size and np.all(np.isfinite(a))
        assert st["refine_iters"] > 0
        rel = float((np.abs(a - b) / (np.abs(b) + m.atol[None, :])).max())
        if all(st[k] == rs[k] for k in ("loop_iters", "bdf_steps", "newton_iters", "nan_repairs")):
            same += 1
            worst_same = max(worst_same, rel)
        print(f"[batch] S={m.S} rel={rel:.2e} refine={st['refine_iters']} "
              f"fallbacks={st['refine_fallbacks']}", flush=True)
    print(f"[batch] {same}/{len(models)} instances on the FP64 decision sequence, "
          f"max rel err among them {worst_same:.2e}", flush=True)
    assert same >= 1
    for dm in dms:
        dm.close()


def test_batch_plan_replays_the_batched_form(gpu_lib):
    """lassie_batch_plan_*: the graph replay gives the batched form's results
    bit for bit, twice in a row, and picks up new rates; the autotuner reports
    its choice."""
    models = synth.uneven_batch(10, 4242, sizes=(16, 48, 100, 200))
    t = synth.t_vector(0, 2, 9)
    kw = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
    print("[plan] 10 uneven instances", flush=True)
    dms = [_lib.DeviceModel(m) for m in models]
    ref, ref_y, ref_st, _, _ = _lib.solve_batch(dms, t, **kw)
    plan = _lib.BatchPlan(dms, t, **kw)
    variant, rt, trial = plan.info()
    assert variant in (0, 1, 2) and rt in (1, 2, 4)
    for _ in range(2):
        outs, ys, st, ms, rc = plan.run()
        assert rc == 0 and ms > 0
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
