"""lassie_solve_batch_static, the batched entry point that allocates no
device memory (its buffers live in one fixed array per GPU), against
lassie_solve_batch through the C ABI: bit-identical trajectories (NaN rows of
capped instances included), final states and counters on two calls in a
row, the second with a smaller input, so contents left in the shared array by
the first call must not leak into the second; batches beyond its limits are
refused with LASSIE_ERR_UNSUPPORTED before any device work."""
import numpy as np
import pytest

from lassie_amd import _lib, synth

pytestmark = pytest.mark.gpu

BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
KEYS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters", "nan_repairs", "lu_failures",
        "samples", "lu_factorizations", "t_final", "dt_final", "newdt_final")


def _both(models, t, **kw):
    dms = [_lib.DeviceModel(m) for m in models]
    try:
        a = _lib.solve_batch(dms, t, raise_on_maxiter=False, **kw)
        b = _lib.solve_batch(dms, t, raise_on_maxiter=False, static=True, **kw)
    finally:
        for dm in dms:
            dm.close()
    return a, b


def _same(a, b):
    assert a[4] == b[4]
    for oa, ob in zip(a[0], b[0]):
        np.testing.assert_array_equal(oa, ob)  # NaN rows at the same places
    for ya, yb in zip(a[1], b[1]):
        np.testing.assert_array_equal(ya, yb)
    for sa, sb in zip(a[2], b[2]):
        for k in KEYS:
            assert sa[k] == sb[k], k


def test_static_batch_equals_the_batched_form_twice(gpu_lib):
    models = synth.uneven_batch(40, 5000, sizes=(16, 32, 64))
    t = synth.t_vector(0.0, 20.0, 41)
    print("[static] 40 instances, S in {16, 32, 64}, Newton cap 6000", flush=True)
    a, b = _both(models, t, max_newton=6000, **BENCH)
    assert a[4] == _lib.LASSIE_ERR_MAXITER  # some instances are capped: their late rows are NaN
    assert any(np.isnan(o).any() for o in b[0]) and any(not np.isnan(o).any() for o in b[0])
    _same(a, b)
    small = [synth.stiff(S, 2 * S, 5200 + i, (1e-4, 1e4)) for i, S in enumerate((7, 16, 33))]
    t2 = synth.t_vector(0.0, 5.0, 11)
    print("[static] second call: 3 smaller instances, fewer samples", flush=True)
    a2, b2 = _both(small, t2, **BENCH)
    assert a2[4] == 0
    assert all(np.isfinite(o).all() for o in b2[0])
    _same(a2, b2)


def test_static_batch_refuses_what_it_cannot_hold(gpu_lib):
    t = synth.t_vector(0.0, 1.0, 3)
    big = _lib.DeviceModel(synth.stiff(65, 130, 1))
    with pytest.raises(_lib.LassieError) as e:
        _lib.solve_batch([big], t, static=True)
    assert e.value.code == _lib.LASSIE_ERR_UNSUPPORTED and "64" in str(e.value)
    big.close()
    m = _lib.DeviceModel(synth.stiff(64, 128, 2))
    with pytest.raises(_lib.LassieError) as e:
        _lib.solve_batch([m], t, static=True, precision=_lib.PREC_FP32_REFINE)
    assert e.value.code == _lib.LASSIE_ERR_UNSUPPORTED
    with pytest.raises(_lib.LassieError) as e:  # 20000 x 64 x 8 B x 64 models of output: beyond 64 MiB
        _lib.solve_batch([m] * 64, synth.t_vector(0.0, 1.0, 20000), static=True)
    assert e.value.code == _lib.LASSIE_ERR_UNSUPPORTED and "bytes" in str(e.value)
    m.close()
