"""Differentiation form on the GPU (lassie_forward_batch /
lassie_adjoint_batch, through the C ABI):
  - the forward pass records the oracle's step sequence (kinds, step sizes,
    sample positions) and trajectories bit for bit;
  - the reverse pass (checkpoint replay every 32 steps, transposed solves,
    sparse vector-Jacobian products) equals the oracle's discrete adjoint,
    itself pinned against forward sensitivities and central differences on
    the CPU, to 1e-9 of the largest gradient component (FP64, different
    summation order);
  - the torch.autograd front end returns the same gradient, within 1e-5 of
    central differences with h = 1e-6 in log k.
"""
import dataclasses

import numpy as np
import pytest

from lassie_amd import _lib, synth
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu
CTRL = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
EPS = 1e-12


def _models():
    ms = [cases.mixed(), cases.with_feed(), synth.stiff(16, 24, 7), synth.stiff(48, 96, 3),
          synth.stiff(130, 260, 5)]
    for m in ms[:2]:
        m.atol = np.full(m.S, 1e-14)
    return ms


def test_forward_records_the_oracle_step_sequence(gpu_lib):
    t = synth.t_vector(0, 1, 6)
    ms = _models()
    print(f"[adjoint] forward record, sizes {[m.S for m in ms]}", flush=True)
    dms = [_lib.DeviceModel(m) for m in ms]
    outs, stats, _, tape = _lib.forward_batch(dms, t, **CTRL)
    for i, m in enumerate(ms):
        o_out, o_st, rc, otape = O.solve_tape(m, t, **CTRL)
        assert rc == 0
        assert np.array_equal(outs[i], o_out), i
        dt, kind, ss = tape.read(i)
        okind, odt = otape.steps()
        assert np.array_equal(kind, okind) and np.array_equal(dt, odt), i
        assert np.array_equal(ss, otape.sample_step), i
        assert stats[i]["bdf_steps"] == o_st["bdf_steps"]
    tape.close()


def test_adjoint_matches_oracle(gpu_lib):
    t = synth.t_vector(0, 1, 6)
    ms = _models()
    print(f"[adjoint] reverse pass, sizes {[m.S for m in ms]}", flush=True)
    dms = [_lib.DeviceModel(m) for m in ms]
    outs, _, _, tape = _lib.forward_batch(dms, t, **CTRL)
    rng = np.random.default_rng(4)
    Ws = [rng.standard_normal(o.shape) for o in outs]
    gks, gxs, ms_dev = tape.adjoint(Ws)
    for i, m in enumerate(ms):
        n, mx = tape.info(i)
        _, _, _, otape = O.solve_tape(m, t, **CTRL)
        ogk, ogx = O.adjoint(m, otape, Ws[i])
        sk = np.abs(ogk).max()
        sx = np.abs(ogx).max()
        print(f"  model {i}: S={m.S} steps={n} max newton={mx} |dgk|/max={np.abs(gks[i] - ogk).max() / sk:.2e}",
              flush=True)
        assert np.abs(gks[i] - ogk).max() <= 1e-9 * sk, i
        assert np.abs(gxs[i] - ogx).max() <= 1e-9 * sx, i
    # a second reverse pass over the same tape gives the same bits
    gks2, _, _ = tape.adjoint(Ws)
    assert all(np.array_equal(a, b) for a, b in zip(gks, gks2))
    tape.close()


def test_rates_changed_after_forward_are_rejected(gpu_lib):
    m = synth.stiff(16, 24, 7)
    dm = _lib.DeviceModel(m)
    outs, _, _, tape = _lib.forward_batch([dm], synth.t_vector(0, 0.5, 3), **CTRL)
    dm.set_rates(m.k * 2)
    with pytest.raises(_lib.LassieError):
        tape.adjoint([np.ones_like(outs[0])])
    tape.close()


def test_step_record_overflow_is_an_error(gpu_lib):
    dm = _lib.DeviceModel(synth.stiff(16, 24, 7))
    with pytest.raises(_lib.LassieError) as e:
        _lib.forward_batch([dm], synth.t_vector(0, 1, 6), max_steps=5, **CTRL)
    assert e.value.code == _lib.LASSIE_ERR_NOMEM


def _targets(ms, hidden, t):
    return np.stack([O.solve(dataclasses.replace(m, k=hidden[b]), t, **CTRL)[0] for b, m in enumerate(ms)])


def test_autograd_log_loss_gradient(gpu_lib):
    torch = pytest.importorskip("torch")
    from lassie_amd import autograd as AG
    ms, hidden = synth.parameterisations(24, 48, 3, seed=2)
    t = synth.t_vector(0, 1, 5)
    print("[adjoint] autograd, 3 parameterisations of 24x48", flush=True)
    target = _targets(ms, hidden, t)
    lk0 = np.log(np.stack([m.k for m in ms]))
    # the library integrates with k = exp(log k), which is not always the same double as k
    ms = [dataclasses.replace(m, k=np.exp(lk0[b])) for b, m in enumerate(ms)]
    dms = [_lib.DeviceModel(m) for m in ms]
    log_k = torch.tensor(lk0, requires_grad=True)
    traj = AG.trajectories(log_k, dms, t, **CTRL)
    loss = AG.log_loss(traj, torch.from_numpy(target), EPS)
    loss.backward()
    g = log_k.grad.numpy()
    h = 1e-6
    for b, m in enumerate(ms):
        out, _, _, otape = O.solve_tape(m, t, **CTRL)
        assert np.array_equal(out, traj[b].detach().numpy())
        gout = 2.0 * (np.log(out + EPS) - np.log(target[b] + EPS)) / (out + EPS)
        ogk, _ = O.adjoint(m, otape, gout)
        scale = np.abs(ogk).max()
        assert np.abs(g[b] - ogk).max() <= 1e-9 * scale

        def loss_b(o):
            return float(np.sum((np.log(o + EPS) - np.log(target[b] + EPS)) ** 2))

        lk = np.log(m.k)
        for r in range(0, m.R, 6):
            e = np.zeros(m.R)
            e[r] = h
            fd = (loss_b(O.replay(m, otape, len(t), k=np.exp(lk + e)))
                  - loss_b(O.replay(m, otape, len(t), k=np.exp(lk - e)))) / (2 * h)
            assert abs(g[b, r] - fd) <= 1e-5 * scale, (b, r, g[b, r], fd)


def test_differentiation_shape_512x1024(gpu_lib):
    """The differentiation workload's network size (512 species x 1024
    reactions, k logU[1e-4, 1e4]) on a short horizon, 4 parameterisations."""
    ms, hidden = synth.parameterisations(512, 1024, 4, seed=0)
    t = synth.t_vector(0, 0.05, 5)
    print("[adjoint] 4 parameterisations of 512x1024, t in [0, 0.05]", flush=True)
    dms = [_lib.DeviceModel(m) for m in ms]
    outs, _, _, tape = _lib.forward_batch(dms, t, **CTRL)
    rng = np.random.default_rng(8)
    Ws = [rng.standard_normal(o.shape) for o in outs]
    gks, gxs, _ = tape.adjoint(Ws)
    for i, m in enumerate(ms[:2]):
        o_out, _, _, otape = O.solve_tape(m, t, **CTRL)
        assert np.array_equal(outs[i], o_out)
        ogk, ogx = O.adjoint(m, otape, Ws[i])
        n, mx = tape.info(i)
        print(f"  parameterisation {i}: steps={n} max newton={mx}", flush=True)
        assert np.abs(gks[i] - ogk).max() <= 1e-9 * np.abs(ogk).max(), i
    tape.close()


def test_nonfinite_gradient_is_reported(gpu_lib):
    """Parameterisations 0 and 108 of the 512 x 1024 workload on t in
    [0, 0.2]: 108 has a BDF1 step whose modified Newton iteration ran all
    10000 iterations without converging, so its gradient is not finite (the
    oracle's exact adjoint overflows too).  The reverse pass returns
    LASSIE_ERR_NONFINITE, lassie_adjoint_status names the model and counts the
    capped Newton steps, and model 0's gradient equals the oracle's."""
    ms, _ = synth.parameterisations(512, 1024, 109, seed=0)
    ms = [ms[0], ms[108]]
    t = synth.t_vector(0, 0.2, 50)
    print("[adjoint] parameterisations 0 and 108 of 512x1024, t in [0, 0.2]", flush=True)
    dms = [_lib.DeviceModel(m) for m in ms]
    outs, _, _, tape = _lib.forward_batch(dms, t, **CTRL)
    rng = np.random.default_rng(2)
    Ws = [rng.standard_normal(o.shape) for o in outs]
    assert tape.status(0)[0] == -1
    with pytest.raises(_lib.LassieError) as err:
        tape.adjoint(Ws)
    assert err.value.code == _lib.LASSIE_ERR_NONFINITE
    gks, gxs, _ = tape.adjoint(Ws, allow_nonfinite=True)
    assert tape.status(0) == (1, 0)
    finite, capped = tape.status(1)
    assert finite == 0 and capped >= 1
    assert np.isfinite(gks[0]).all() and not np.isfinite(gks[1]).all()
    o_out, _, _, otape = O.solve_tape(ms[0], t, **CTRL)
    ogk, _ = O.adjoint(ms[0], otape, Ws[0])
    assert np.abs(gks[0] - ogk).max() <= 1e-9 * np.abs(ogk).max()
    tape.close()
