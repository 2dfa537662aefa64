"""Wave-level batched kernel (one wavefront per instance, S <= 64) against
the oracle, bit for bit: trajectories, final states and every controller
counter.  Covers ragged sizes 1..64, the hand-built models (every RHS and
Jacobian branch, a singular Newton matrix, feed species), both controllers
(the reference's and the bench's), the per-solve Newton-iteration cap, the
equality of the wave-level and workgroup-level kernels on the same instances,
and plans that mix the two kernels.
"""
import numpy as np
import pytest

from lassie_amd import _lib, synth
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
KEYS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs",
        "lu_failures", "samples", "lu_factorizations")


def _check(models, t, outs, ys, stats, **kw):
    for m, out, y, st in zip(models, outs, ys, stats):
        o_out, o_st, o_rc, _, o_y = O.solve(m, t, **kw)
        for key in KEYS:
            if key == "lu_factorizations":
                continue
            assert st[key] == o_st[key], (m.S, key, st[key], o_st[key])
        n = int(o_st["samples"])
        np.testing.assert_array_equal(out[:n], o_out[:n])
        np.testing.assert_array_equal(y, o_y)


def _run(models, t, **kw):
    print(f"[wave] sizes {[m.S for m in models]} kw {kw}", flush=True)
    dms = [_lib.DeviceModel(m) for m in models]
    try:
        outs, ys, stats, ms, rc = _lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
    finally:
        for dm in dms:
            dm.close()
    return outs, ys, stats, rc


@pytest.mark.parametrize("ctrl", ["bench", "reference"])
def test_wave_ragged_sizes_bit_exact(gpu_lib, ctrl):
    sizes = (2, 3, 7, 15, 16, 17, 24, 31, 32, 33, 47, 48, 63, 64)
    models = [synth.stiff(S, 2 * S, 5100 + i, (1e-4, 1e4)) for i, S in enumerate(sizes)]
    t = synth.t_vector(0, 20, 41)
    kw = dict(BENCH, max_iters=200_000) if ctrl == "bench" else dict(max_iters=20_000)
    outs, ys, stats, rc = _run(models, t, **kw)
    _check(models, t, outs, ys, stats, **kw)


def test_wave_hand_models_bit_exact(gpu_lib):
    models = [cases.decay(), cases.reversible(), cases.chain(), cases.dimer(), cases.mixed(),
              cases.autocatalytic_singular(), cases.with_feed(), cases.hub(20)]
    t = synth.t_vector(0, 5, 11)
    for kw in (dict(BENCH, max_iters=200_000), dict(max_iters=50_000)):
        outs, ys, stats, rc = _run(models, t, **kw)
        _check(models, t, outs, ys, stats, **kw)
    # the singular model fails its first LU and shrinks the BDF step
    _, _, stats, _ = _run([cases.autocatalytic_singular()], t, max_iters=50_000)
    assert stats[0]["lu_failures"] > 0


def test_wave_c4_generator_instances_bit_exact(gpu_lib):
    """Instances of the 2048-instance uneven workload's generator (S in
    {16, 32, 64}, R = 2S, k logU[1e-4, 1e4]) on its horizon [0, 50] with
    100 samples, under the bench controller."""
    models = synth.uneven_batch(24, 1000, sizes=(16, 32, 64))
    t = synth.t_vector(0.0, 50.0, 100)
    kw = dict(BENCH, max_newton=300_000)
    outs, ys, stats, rc = _run(models, t, **kw)
    _check(models, t, outs, ys, stats, **kw)


def test_wave_equals_workgroup_kernel(gpu_lib):
    """opts.reserved bit 2 sends every instance to the workgroup-level
    kernel: both kernels give the same bits and the same counters."""
    models = synth.uneven_batch(12, 777, sizes=(5, 16, 32, 50, 64))
    t = synth.t_vector(0, 10, 21)
    kw = dict(BENCH, max_iters=200_000)
    a = _run(models, t, **kw)
    b = _run(models, t, reserved=2, **kw)
    for oa, ob in zip(a[0], b[0]):
        np.testing.assert_array_equal(oa, ob)
    for ya, yb in zip(a[1], b[1]):
        np.testing.assert_array_equal(ya, yb)
    for sa, sb in zip(a[2], b[2]):
        for key in KEYS:
            assert sa[key] == sb[key], key


def test_wave_newton_cap_matches_oracle(gpu_lib):
    """max_newton stops an instance where the oracle stops with the same cap;
    the samples stored before it and the state at that point agree, and the
    batch reports LASSIE_ERR_MAXITER."""
    models = [synth.stiff(S, 2 * S, 6100 + i, (1e-4, 1e4)) for i, S in enumerate((16, 40, 64))]
    t = synth.t_vector(0.0, 50.0, 100)
    kw = dict(BENCH, max_newton=500)
    outs, ys, stats, rc = _run(models, t, **kw)
    assert rc == _lib.LASSIE_ERR_MAXITER
    for m, st in zip(models, stats):
        assert 500 <= st["newton_iters"] < 500 + 10000 and st["samples"] < t.size
    _check(models, t, outs, ys, stats, **kw)


def test_plan_mixes_wave_and_workgroup_kernels(gpu_lib):
    """A plan over small and large instances runs both kernels, reports how
    it split them and their times, and equals lassie_solve_batch."""
    models = synth.uneven_batch(14, 4343, sizes=(16, 32, 64, 128, 200))
    t = synth.t_vector(0, 2, 9)
    kw = dict(BENCH)
    dms = [_lib.DeviceModel(m) for m in models]
    ref, ref_y, ref_st, _, _ = _lib.solve_batch(dms, t, **kw)
    plan = _lib.BatchPlan(dms, t, **kw)
    n_small = sum(m.S <= 64 for m in models)
    for _ in range(2):
        outs, ys, st, ms, rc = plan.run()
        n_wave, n_wg, kms = plan.timing()
        assert (n_wave, n_wg) == (n_small, len(models) - n_small)
        assert kms[0] > 0 and kms[1] > 0 and plan.progress() == len(models)
        for a, b in zip(outs, ref):
            assert np.array_equal(a, b)
        for a, b in zip(ys, ref_y):
            assert np.array_equal(a, b)
    plan.close()
    for dm in dms:
        dm.close()
